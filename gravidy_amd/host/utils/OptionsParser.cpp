// OptionsParser.cpp -- getopt_long front end with the reference's flags and
// defaults (OptionsParser.cpp:41-89, 111-423): -i input, -t time (float,
// default 1), -r resume from an .info file and its last text snapshot,
// -o output, -s softening (float, default 1e-4), -e eta (float, default
// 0.01), -p screen, -z interval (float, default 0.125), -a all, -g gpus.
// Lagrange radii (-l) are not part of this build and are rejected with a
// message; --checkpoint and --restart give a bit-exact binary restart
// beside the (lossy) text resume.
#include "OptionsParser.hpp"

#include <getopt.h>
#include <sys/stat.h>

#include <algorithm>
#include <map>
#include <set>

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
          "  -r [ --resume ] <filename>.info  Resume a simulation with an .info file\n"
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

// Digits of the snapshot counter (OptionsParser.cpp:393-401).
static unsigned int snapshot_digits(float integration_time, float interval_time) {
    unsigned int len = 0;
    unsigned int tmp = (unsigned int)(integration_time / interval_time);
    do {
        ++len;
        tmp /= 10;
    } while (tmp);
    return len + 1;
}

// Resume from a run's .info file (OptionsParser.cpp:122-324): "# Key: value"
// lines, '#' and blanks dropped from keys and blanks from values; any key
// outside the eight the info writer emits is an error.  -t must exceed the
// file's IntegrationTime; softening and eta come from the file unless -s /
// -e give positive values; InputFilename names the output (".out", -o is
// not consulted, as there); the run restarts from OutputFilename's snapshot
// number SnapshotNumber, whose "# Time:" header gives the start time
// (NbodySystem.cpp:106-146).  The reference names that snapshot with four
// digits (OptionsParser.cpp:305-306), while its writer pads to the width of
// the run's output count (Logger.cpp:80), so it finds only snapshots of
// runs with 100-999 outputs; here the four-digit name is tried first, then
// every other padding.  The reference leaves the output interval and the
// snapshot-name width unset when resuming; here they follow -z (default
// 0.125) and -t as for a new run.
bool OptionsParser::check_resume() {
    std::cerr << "Reading resume_filename: " << resume_filename << std::endl;
    std::ifstream rfile(resume_filename.c_str());
    std::map<std::string, std::string> param;
    std::string key, value;
    while (rfile.good() && std::getline(rfile, key, ':')) {
        key.erase(std::remove(key.begin(), key.end(), '#'), key.end());
        key.erase(std::remove(key.begin(), key.end(), ' '), key.end());
        if (!std::getline(rfile, value)) break;
        value.erase(std::remove(value.begin(), value.end(), ' '), value.end());
        param[key] = value;
    }
    static const char *const allowed[] = {"NumberParticles", "Softening",      "EtaTimestep",
                                          "IntegrationTime", "PrintScreen",    "InputFilename",
                                          "OutputFilename",  "SnapshotNumber"};
    const std::set<std::string> ok(allowed, allowed + 8);
    for (std::map<std::string, std::string>::const_iterator it = param.begin(); it != param.end();
         ++it) {
        if (!ok.count(it->first)) {
            std::cerr << "Invalid param" << std::endl
                      << "gravidy: Invalid option -- " << it->first << " in the info file"
                      << std::endl;
            usage(std::cerr);
            return false;
        }
        std::cerr << it->first << " -> " << it->second << std::endl;
    }
    const double config_time = std::strtod(param["IntegrationTime"].c_str(), NULL);
    if (config_time >= integration_time) {
        std::cerr << "gravidy: option requires an argument -- 'time' greater than the one "
                     "specifid in the info file"
                  << std::endl;
        usage(std::cerr);
        return false;
    }
    snapshot_time = (float)config_time;
    const float file_soft = (float)std::strtod(param["Softening"].c_str(), NULL);
    if (!have_soft || !(softening > 0.0f)) {
        if (have_soft) std::cerr << "gravidy: softening cannot be negative" << std::endl;
        softening = file_soft;
    }
    const float file_eta = (float)std::strtod(param["EtaTimestep"].c_str(), NULL);
    if (!have_eta || !(eta > 0.0f)) {
        if (have_eta) std::cerr << "gravidy: eta cannot be negative" << std::endl;
        eta = file_eta;
    }
    ops.print_screen = std::atoi(param["PrintScreen"].c_str()) != 0;
    input_filename = param["InputFilename"];
    output_filename = input_filename + ".out";
    snapshot_number = (unsigned int)std::atoi(param["SnapshotNumber"].c_str());
    const std::string base = param["OutputFilename"] + ".snapshot_";
    std::vector<int> widths(1, 4);
    for (int w = 1; w <= 10; ++w)
        if (w != 4) widths.push_back(w);
    for (size_t k = 0; k < widths.size() && snapshot_filename.empty(); ++k) {
        std::ostringstream s;
        s << std::setw(widths[k]) << std::setfill('0') << snapshot_number;
        if (file_exists(base + s.str())) snapshot_filename = base + s.str();
    }
    if (snapshot_filename.empty()) {
        std::ostringstream s;
        s << std::setw(4) << std::setfill('0') << snapshot_number;
        std::cerr << "gravidy: cannot access " << base << s.str()
                  << ": No such file or directory\n(Check the OutputFilename and SnapshotNumber)"
                  << std::endl;
        return false;
    }
    return true;
}

bool OptionsParser::check_options() {
    if (!parse_ok || help) {
        usage(std::cerr);
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
    if (have_resume) {
        if (!file_exists(resume_filename)) {
            std::cerr << "gravidy: cannot access " << resume_filename
                      << ": No such file or directory" << std::endl;
            return false;
        }
        if (!restart_filename.empty()) {
            std::cerr << "gravidy: -r (text resume) and --restart (binary checkpoint) exclude "
                         "each other"
                      << std::endl;
            return false;
        }
        if (!check_resume()) return false;
        resume = true;
    } else {
        if (!have_input) {
            std::cerr << "gravidy: option requires an argument -- 'input'" << std::endl;
            usage(std::cerr);
            return false;
        }
        input_filename = opt_input;
        if (!file_exists(input_filename)) {
            std::cerr << "gravidy: cannot access " << input_filename
                      << ": No such file or directory" << std::endl;
            return false;
        }
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
    length_output_number = snapshot_digits(integration_time, interval_time);
    if (resume) return true;
    output_filename = (have_output ? opt_output : input_filename) + ".out";
    snapshot_number = 0;
    return true;
}
