// OptionsParser.hpp -- command-line options of gravidy-hip: the reference's
// flags (OptionsParser.cpp:41-89) parsed with getopt_long, plus the
// backend's --precision, --dt-min, --dt-max, --checkpoint and --restart.
#ifndef H4_HOST_OPTIONSPARSER_HPP
#define H4_HOST_OPTIONSPARSER_HPP

#include "../common.hpp"

class OptionsParser {
  public:
    OptionsParser(int argc, char *argv[]);
    ~OptionsParser();

    std::string input_filename;
    std::string output_filename;
    std::string resume_filename;
    std::string snapshot_filename;
    float integration_time;
    float interval_time;
    int gpus;
    float snapshot_time;
    float softening;
    float eta;
    unsigned int snapshot_number;
    unsigned int length_output_number;
    bool resume;
    options ops;
    int precision;
    double dt_min;
    double dt_max;
    std::string checkpoint_filename;  // --checkpoint: binary checkpoint at outputs
    std::string restart_filename;     // --restart: continue from a checkpoint

    bool file_exists(std::string filename);
    bool check_options();
    bool check_resume();
    void usage(std::ostream &os) const;

  private:
    bool parse_ok;
    bool help;
    bool have_input, have_output, have_soft, have_eta, have_interval, have_resume;
    std::string opt_input, opt_output, opt_precision;
};

#endif
