// NbodySystem.cpp -- state construction and the initial-condition reader.
#include "NbodySystem.hpp"

#include <algorithm>
#include <cctype>

// Softening squared is formed from the float softening in float, exactly as
// the reference does (NbodySystem.cpp:50); eta stays float.
NbodySystem::NbodySystem(OptionsParser op) {
    input_filename = op.input_filename;
    output_filename = op.output_filename;
    resume_filename = op.resume_filename;
    snapshot_filename = op.snapshot_filename;
    integration_time = op.integration_time;
    interval_time = op.interval_time;
    snapshot_time = op.snapshot_time;
    snapshot_number = op.snapshot_number;
    length_snap = op.length_output_number;
    float soft2 = op.softening * op.softening;
    e2 = soft2;
    eta = op.eta;
    ops = op.ops;
    iterations = 0;
    gpus = op.gpus;
    resume = op.resume;
    precision = op.precision;
    dt_min = op.dt_min;
    dt_max = op.dt_max;
    en.ini = en.end = en.tmp = en.kinetic = en.potential = 0.0;
    std::memset(&gtime, 0, sizeof(gtime));
    m_g = 0.0;
    max_mass = 0.0f;
    total_mass = 0.0f;
    n = 0;
    h_id = NULL;
    h_r = NULL;
    h_v = NULL;
}

NbodySystem::~NbodySystem() { free_base_attributes(); }

// Reads "id m x y z vx vy vz" lines (NbodySystem.cpp:79-207 semantics):
// lines whose first non-blank character is '#' are comments; until a start
// time has been found, a comment's "time <value>" pair sets it; mass is
// parsed as float; a data line with other than 8 columns stops the read.
// Like the reference's getline loop, a last line without a trailing newline
// is not read.
bool NbodySystem::read_input_file() {
    total_mass = 0;
    max_mass = 0.0;
    bool got_time = false;
    std::ifstream file((resume ? snapshot_filename : input_filename).c_str());
    if (!file.is_open()) return false;
    std::string line;
    bool ok = true;
    std::getline(file, line);
    while (file.good()) {
        std::string::size_type first = line.find_first_not_of(" ");
        if (first != std::string::npos && line[first] == '#') {
            if (!got_time) {
                std::string clean;
                for (size_t k = 0; k < line.size(); ++k)
                    if (line[k] != '#' && line[k] != ':')
                        clean += (char)std::tolower((unsigned char)line[k]);
                std::istringstream iss(clean);
                std::vector<std::string> tok;
                std::string w;
                while (iss >> w) tok.push_back(w);
                for (size_t t = 0; t + 1 < tok.size(); ++t) {
                    if (tok[t] == "time") {
                        snapshot_time = std::strtod(tok[t + 1].c_str(), NULL);
                        got_time = true;
                        break;
                    }
                }
            }
        } else {
            std::istringstream iss(line);
            std::vector<std::string> tok;
            std::string w;
            while (iss >> w) tok.push_back(w);
            if (tok.size() != 8) {
                std::cerr << "gravidy: wrong line format in" << input_filename
                          << ": only eight numbers per line are allowed" << std::endl;
                ok = false;
                break;
            }
            file_data tmp;
            tmp.id = (unsigned int)std::atoi(tok[0].c_str());
            tmp.m = (float)std::strtod(tok[1].c_str(), NULL);
            for (int c = 0; c < 3; ++c) {
                tmp.r[c] = std::strtod(tok[2 + c].c_str(), NULL);
                tmp.v[c] = std::strtod(tok[5 + c].c_str(), NULL);
            }
            if (tmp.m > max_mass) max_mass = tmp.m;
            reader.push_back(tmp);
            total_mass += tmp.m;
        }
        std::getline(file, line);
    }
    n = (unsigned int)reader.size();
    return ok;
}

void NbodySystem::alloc_base_attributes(int) {
    h_id = new unsigned int[n];
    h_r = new double4[n];
    h_v = new double4[n];
}

void NbodySystem::free_base_attributes() {
    delete[] h_id;
    delete[] h_r;
    delete[] h_v;
    h_id = NULL;
    h_r = NULL;
    h_v = NULL;
}

void NbodySystem::copy_input_data() {
    for (unsigned int i = 0; i < reader.size(); i++) {
        h_id[i] = reader[i].id;
        h_r[i].x = reader[i].r[0];
        h_r[i].y = reader[i].r[1];
        h_r[i].z = reader[i].r[2];
        h_r[i].w = reader[i].m;
        h_v[i].x = reader[i].v[0];
        h_v[i].y = reader[i].v[1];
        h_v[i].z = reader[i].v[2];
        h_v[i].w = 0.0;
    }
    std::vector<file_data>().swap(reader);
}
