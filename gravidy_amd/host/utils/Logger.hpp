// Logger.hpp -- output layer: energy log, info file, snapshots, messages.
// Method names and file formats are the reference's (Logger.cpp:77-460) so
// runs of either backend produce comparable files.
#ifndef H4_HOST_LOGGER_HPP
#define H4_HOST_LOGGER_HPP

#include "../NbodySystem.hpp"
#include "../common.hpp"

class Logger {
  public:
    Logger(NbodySystem *ns);
    ~Logger();

    NbodySystem *ns;
    bool print_screen;
    std::ofstream out_file;
    std::ofstream info_file;
    std::ostream *gstream;
    std::string ofname;
    std::string ofname_info;

    std::string get_timestamp();
    void print_info();
    void add_info(std::string key, std::string value);
    void write_info();
    void write_snapshot(unsigned int snapshot_number, double ITIME);
    void print_all(double ITIME, unsigned int snapshot_number);
    void print_energy_log(double ITIME, unsigned int iterations, long long interactions,
                          unsigned int nsteps, double new_energy);
    void print_lagrange_radii(double ITIME, std::vector<double> lagrange_radii);

    void log(int type, std::string msg);
    void log_error(std::string msg);
    void log_warning(std::string msg);
    void log_success(std::string msg);
    void log_info(std::string msg);
};

#endif
