// Logger.cpp -- the reference's output formats, written for this framework.
//
// Stream-state detail that the formats depend on: the reference routes
// precision changes of the snapshot writer through `gstream`, which points
// at `out_file` in file mode and at std::cout in screen mode (-p).  In file
// mode snapshots therefore come out in scientific notation with 6 digits,
// in screen mode with 4 (Logger.cpp:84-97).  The same holds here.
#include "Logger.hpp"

#include <sys/stat.h>

#include <cctype>
#include <ctime>

static const char *const kColor[] = {"\033[0;31m", "\033[0;33m", "\033[0;32m", "\033[0;36m"};
static const char *const kLevel[] = {" [ERROR] ", " [WARNING] ", " [SUCCESS] ", " [INFO] "};
static const char *const kBlue = "\033[0;34m";
static const char *const kReset = "\033[m";

Logger::Logger(NbodySystem *ns) : ns(ns), gstream(&std::cout) {
    print_screen = ns->ops.print_screen;
    ofname = ns->output_filename;
    ofname_info = ns->resume ? ns->resume_filename : ofname + ".info";
}

Logger::~Logger() {}

std::string Logger::get_timestamp() {
    time_t now = time(0);
    tm *t = localtime(&now);
    std::ostringstream s;
    s << "[" << std::setw(4) << std::setfill('0') << 1900 + t->tm_year << "-" << std::setw(2)
      << 1 + t->tm_mon << "-" << std::setw(2) << t->tm_mday << " " << std::setw(2)
      << 1 + t->tm_hour << ":" << std::setw(2) << 1 + t->tm_min << ":" << std::setw(2)
      << 1 + t->tm_sec << "]";
    return s.str();
}

// Snapshot: "# Time:" in fixed notation with 4 decimals, then per particle
// id (width 6) and m, x, y, z, vx, vy, vz (width 15, scientific).
void Logger::write_snapshot(unsigned int snapshot_number, double ITIME) {
    std::ostringstream s;
    s << std::setw(ns->length_snap) << std::setfill('0') << snapshot_number;
    std::string name = ofname + ".snapshot_" + s.str();
    out_file.open(name.c_str(), std::ios::out);
    out_file << std::fixed;
    out_file.precision(4);
    out_file << "# Time:" << std::setw(6) << ITIME << std::endl;
    const std::streamsize prec = print_screen ? 4 : 6;
    for (unsigned int i = 0; i < ns->n; i++) {
        out_file << std::setw(6) << std::right << ns->h_id[i];
        out_file << std::scientific;
        out_file.precision(prec);
        const double row[7] = {ns->h_r[i].w, ns->h_r[i].x, ns->h_r[i].y, ns->h_r[i].z,
                               ns->h_v[i].x, ns->h_v[i].y, ns->h_v[i].z};
        for (int k = 0; k < 7; ++k) out_file << std::setw(15) << std::right << row[k];
        out_file << std::endl;
    }
    out_file.close();
}

void Logger::add_info(std::string key, std::string value) {
    info_file.open(ofname_info.c_str(), std::ios_base::app);
    info_file << std::setw(2) << std::left << "#" << std::setw(20) << std::left << key
              << std::setw(20) << std::right << value << std::endl;
    info_file.close();
}

static bool exists(const std::string &f) {
    struct stat b;
    return stat(f.c_str(), &b) != -1;
}

// Moves an existing info file aside (".old", then ".old<k>") before
// writing the run's keys (Logger.cpp:136-217).
void Logger::write_info() {
    if (exists(ofname_info)) {
        std::string last = ofname_info.substr(ofname_info.size() - 1);
        std::string newname = ofname_info + ".old";
        int ext = 0;
        while (exists(newname)) {
            if (std::isdigit((unsigned char)last[0])) ext = std::atoi(last.c_str()) + 1;
            newname = ofname_info + ".old" + SSTR(ext);
            ext++;
        }
        if (std::rename(ofname_info.c_str(), newname.c_str()))
            std::cerr << "gravidy: Failed to rename info file from " << ofname_info << " to "
                      << newname << std::endl;
    }
    info_file.open(ofname_info.c_str(), std::ios::out);
    info_file << std::setw(2) << std::left << "#" << std::setw(20) << std::left
              << "NumberParticles:" << std::setw(20) << std::right << ns->n << std::endl;
    info_file << std::setw(2) << std::left << "#" << std::setw(20) << std::left << "Softening:"
              << std::setw(20) << std::right << std::sqrt(ns->e2) << std::endl;
    info_file << std::setw(2) << std::left << "#" << std::setw(20) << std::left
              << "EtaTimestep:" << std::setw(20) << std::right << ns->eta << std::endl;
    info_file << std::setw(2) << std::left << "#" << std::setw(20) << std::left
              << "IntegrationTime:" << std::setw(20) << std::right << ns->integration_time
              << std::endl;
    info_file << std::setw(2) << std::left << "#" << std::setw(20) << std::left
              << "PrintScreen:" << std::setw(20) << std::right << print_screen << std::endl;
    info_file << std::setw(2) << std::left << "#" << std::setw(20) << std::left
              << "InputFilename:" << std::setw(20) << std::right << ns->input_filename
              << std::endl;
    info_file << std::setw(2) << std::left << "#" << std::setw(20) << std::left
              << "OutputFilename:" << std::setw(20) << std::left << ns->output_filename
              << std::endl;
    info_file.close();
}

// Creates (truncates) the .log file in file mode.
void Logger::print_info() {
    if (print_screen) {
        gstream = &std::cout;
        return;
    }
    std::string name = ofname + ".log";
    out_file.open(name.c_str(), std::ios::out);
    gstream = &out_file;
    out_file.close();
}

void Logger::print_lagrange_radii(double, std::vector<double>) {
    log_error("Lagrange radii are not part of this build");
}

// One line per particle: time, index, m, r, v, a, a1, dt (Logger.cpp:270-322).
void Logger::print_all(double ITIME, unsigned int snapshot_number) {
    std::ostream *os = &std::cout;
    if (!print_screen) {
        std::ostringstream s;
        s << std::setw(4) << std::setfill('0') << snapshot_number;
        std::string name = ofname + ".all.snap_" + s.str();
        out_file.open(name.c_str(), std::ios::out);
        os = &out_file;
    }
    for (unsigned int i = 0; i < ns->n; i++) {
        *os << std::fixed;
        os->precision(4);
        *os << std::setw(6) << ITIME;
        os->precision(2);
        *os << std::setw(6) << i;
        *os << std::scientific;
        os->precision(6);
        const double row[14] = {ns->h_r[i].w, ns->h_r[i].x, ns->h_r[i].y, ns->h_r[i].z,
                                ns->h_v[i].x, ns->h_v[i].y, ns->h_v[i].z, ns->h_f[i].a[0],
                                ns->h_f[i].a[1], ns->h_f[i].a[2], ns->h_f[i].a1[0],
                                ns->h_f[i].a1[1], ns->h_f[i].a1[2], ns->h_dt[i]};
        for (int k = 0; k < 14; ++k) *os << std::setw(15) << std::right << row[k];
        *os << std::endl;
    }
    if (!print_screen) out_file.close();
}

// Energy table (Logger.cpp:324-391): RelE against the previous output,
// CumE against the initial energy; GFLOPS from the interaction count and
// the accumulated update time.
void Logger::print_energy_log(double ITIME, unsigned int iterations, long long interactions,
                              unsigned int nsteps, double new_energy) {
    ns->en.end = new_energy;
    double rel_error = std::abs((ns->en.end - ns->en.tmp) / ns->en.tmp);
    double cum_error = std::abs((ns->en.end - ns->en.ini) / ns->en.ini);
    ns->en.tmp = ns->en.end;
    float time = omp_get_wtime() - ns->gtime.integration_ini;
    if (ITIME != 0.0) ns->gtime.gflops = KERNEL_GFLOP * (interactions / ns->gtime.update_end);

    std::ostream *os = &std::cout;
    if (!print_screen) {
        std::string name = ofname + ".log";
        out_file.open(name.c_str(), std::ios::out | std::ios::app);
        os = &out_file;
    }
    gstream = os;
    if (ITIME == 0.0) {
        *os << std::setw(2) << std::left << "#";
        *os << std::setw(10) << std::right << "IteTime";
        *os << std::setw(15) << std::right << "Iter";
        *os << std::setw(15) << std::right << "Nsteps";
        *os << std::setw(20) << std::right << "Energy";
        *os << std::setw(15) << std::right << "RelE";
        *os << std::setw(15) << std::right << "CumE";
        *os << std::setw(16) << std::right << "ElapsedTime";
        *os << std::setw(12) << std::right << "GFLOPS";
        *os << std::endl;
    }
    *os << std::fixed;
    *os << std::setw(2) << std::right << "00";
    os->precision(3);
    *os << std::setw(10) << std::right << ITIME;
    *os << std::setw(15) << std::right << iterations;
    *os << std::setw(15) << std::right << nsteps;
    *os << std::scientific;
    os->precision(10);
    *os << std::setw(20) << std::right << ns->en.end;
    os->precision(6);
    *os << std::setw(15) << std::right << rel_error;
    *os << std::setw(15) << std::right << cum_error;
    *os << std::setw(16) << std::right << time;
    *os << std::fixed;
    os->precision(3);
    *os << std::setw(12) << std::right << ns->gtime.gflops;
    *os << std::endl;
    if (!print_screen) out_file.close();
}

void Logger::log(int type, std::string msg) {
    if (type < 0 || type > 3) return;
    std::cerr << kBlue << get_timestamp() << kColor[type] << kLevel[type] << kReset << msg
              << std::endl;
}

void Logger::log_error(std::string msg) { log(0, msg); }
void Logger::log_warning(std::string msg) { log(1, msg); }
void Logger::log_success(std::string msg) { log(2, msg); }
void Logger::log_info(std::string msg) { log(3, msg); }
