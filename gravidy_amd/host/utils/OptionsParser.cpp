// OptionsParser.cpp -- getopt_long front end with the reference's flags and
// defaults (OptionsParser.cpp:41-89, 111-423): -i input, -t time (float,
// default 1), -o output, -s softening (float, default 1e-4), -e eta (float,
// default 0.01), -p screen, -z interval (float, default 0.125), -a all,
// -g gpus.  Resume from a text snapshot (-r) and Lagrange radii (-l) are not
// part of this build and are rejected with a message; --checkpoint and
// --restart give a bit-exact binary restart instead.
#include "OptionsParser.hpp"

#include <getopt.h>
#include <sys/stat.h>

#include "h4hip.h"

OptionsParser::OptionsParser(int argc, char *argv[]) {
    integration_time = 1.0f;
    interval_time = (float)D_TIME_MAX;
    gpus = 0;
    snapshot_time = 0.0f;
    softening = (float)E;
    eta = ETA_N;
    snapshot_number = 0;
    length_output_number = 0;
    resume = false;
    ops.print_screen = ops.print_all = ops.print_lagrange = false;
    precision = H4H_FP64;
    dt_min = D_TIME_MIN;
    dt_max = D_TIME_MAX;
    parse_ok = true;
    help = have_input = have_output = have_soft = have_eta = have_interval = have_resume = false;

    static struct option longopts[] = {
        {"help", no_argument, 0, 'h'},      {"input", required_argument, 0, 'i'},
        {"time", required_argument, 0, 't'}, {"resume", required_argument, 0, 'r'},
        {"output", required_argument, 0, 'o'}, {"softening", required_argument, 0, 's'},
        {"eta", required_argument, 0, 'e'},  {"screen", no_argument, 0, 'p'},
        {"interval", required_argument, 0, 'z'}, {"lagrange", no_argument, 0, 'l'},
        {"all", no_argument, 0, 'a'},        {"gpu", required_argument, 0, 'g'},
        {"precision", required_argument, 0, 1}, {"dt-min", required_argument, 0, 2},
        {"dt-max", required_argument, 0, 3}, {"checkpoint", required_argument, 0, 4},
        {"restart", required_argument, 0, 5}, {0, 0, 0, 0}};
    optind = 1;
    int c;
    while ((c = getopt_long(argc, argv, "hi:t:r:o:s:e:pz:lag:", longopts, NULL)) != -1) {
        switch (c) {
            case 'h': help = true; break;
            case 'i': have_input = true; opt_input = optarg; break;
            case 't': integration_time = std::strtof(optarg, NULL); break;
            case 'r': have_resume = true; resume_filename = optarg; break;
            case 'o': have_output = true; opt_output = optarg; break;
            case 's': have_soft = true; softening = std::strtof(optarg, NULL); break;
            case 'e': have_eta = true; eta = std::strtof(optarg, NULL); break;
            case 'p': ops.print_screen = true; break;
            case 'z': have_interval = true; interval_time = std::strtof(optarg, NULL); break;
            case 'l': ops.print_lagrange = true; break;
            case 'a': ops.print_all = true; break;
            case 'g': gpus = std::atoi(optarg); break;
            case 1: opt_precision = optarg; break;
            case 2: dt_min = std::strtod(optarg, NULL); break;
            case 3: dt_max = std::strtod(optarg, NULL); break;
            case 4: checkpoint_filename = optarg; break;
            case 5: restart_filename = optarg; break;
            default: parse_ok = false; break;
        }
    }
}

OptionsParser::~OptionsParser() {}

bool OptionsParser::file_exists(std::string filename) {
    struct stat buffer;
    return stat(filename.c_str(), &buffer) != -1;
}

void OptionsParser::usage(std::ostream &os) const {
    os << "GraviDy (HIP backend)\n"
          "  -h [ --help ]                 Display this message\n"
          "  -i [ --input ] <filename>     Input data filename\n"
          "  -t [ --time ] <value> (=1)    Integration time (In N-body units)\n"
          "  -o [ --output ] <filename>    Output data filename\n"
          "  -s [ --softening ] <value>    Softening parameter (default 1e-4)\n"
          "  -e [ --eta ] <value>          ETA of time-step calculation (default 0.01)\n"
          "  -p [ --screen ]               Print summary in the screen instead of a file\n"
          "  -z [ --interval ] <value>     Output time interval (default 0.125)\n"
          "  -a [ --all ]                  Print all the information of N-particles in every "
          "integration time\n"
          "  -g [ --gpu ] <value> (=0)     GPUs to use (this build: one)\n"
          "  --precision fp64|exact|mixed  Pair-kernel precision (default fp64)\n"
          "  --dt-min <value>              Smallest block timestep (default 2^-23)\n"
          "  --dt-max <value>              Largest block timestep (default 2^-3)\n"
          "  --checkpoint <filename>       Binary checkpoint at every output time and at the end\n"
          "  --restart <filename>          Continue bit for bit from a binary checkpoint\n";
}

bool OptionsParser::check_options() {
    if (!parse_ok || help) {
        usage(std::cerr);
        return false;
    }
    if (have_resume) {
        std::cerr << "gravidy: resuming from an .info file is not part of this build" << std::endl;
        return false;
    }
    if (ops.print_lagrange) {
        std::cerr << "gravidy: Lagrange radii are not part of this build" << std::endl;
        return false;
    }
    if (gpus > 1) {
        std::cerr << "gravidy: this build runs on one GPU" << std::endl;
        return false;
    }
    if (!have_input) {
        std::cerr << "gravidy: option requires an argument -- 'input'" << std::endl;
        usage(std::cerr);
        return false;
    }
    input_filename = opt_input;
    if (!file_exists(input_filename)) {
        std::cerr << "gravidy: cannot access " << input_filename << ": No such file or directory"
                  << std::endl;
        return false;
    }
    if (!opt_precision.empty()) {
        if (opt_precision == "fp64") precision = H4H_FP64;
        else if (opt_precision == "exact") precision = H4H_FP64_EXACT;
        else if (opt_precision == "mixed") precision = H4H_MIXED;
        else {
            std::cerr << "gravidy: unknown precision " << opt_precision << std::endl;
            return false;
        }
    }
    if (!(dt_min > 0.0) || !(dt_max >= dt_min)) {
        std::cerr << "gravidy: need 0 < dt-min <= dt-max" << std::endl;
        return false;
    }
    if (!restart_filename.empty() && !file_exists(restart_filename)) {
        std::cerr << "gravidy: cannot access " << restart_filename << ": No such file or directory"
                  << std::endl;
        return false;
    }
    if (!(interval_time > 0.0f)) {
        std::cerr << "gravidy: interval must be positive" << std::endl;
        return false;
    }
    // Digits of the snapshot counter (OptionsParser.cpp:393-401).
    length_output_number = 0;
    unsigned int tmp = (unsigned int)(integration_time / interval_time);
    do {
        ++length_output_number;
        tmp /= 10;
    } while (tmp);
    length_output_number += 1;
    output_filename = (have_output ? opt_output : input_filename) + ".out";
    snapshot_number = 0;
    resume = false;
    return true;
}
