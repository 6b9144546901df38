// H4Checkpoint.cpp -- binary checkpoint reader and writer (format in
// H4Checkpoint.hpp).
#include "H4Checkpoint.hpp"

#include <cstdio>
#include <cstring>

namespace {

const char MAGIC[8] = {'H', '4', 'H', 'C', 'K', 'P', 'T', 1};

// CRC-32 (ISO-HDLC: reflected polynomial 0xEDB88320, initial and final
// value 0xFFFFFFFF -- the zlib/PNG checksum), table driven.
struct Crc32 {
    uint32_t table[256];
    uint32_t c;
    Crc32() : c(0xFFFFFFFFu) {
        for (uint32_t k = 0; k < 256; ++k) {
            uint32_t x = k;
            for (int b = 0; b < 8; ++b) x = (x & 1u) ? (x >> 1) ^ 0xEDB88320u : x >> 1;
            table[k] = x;
        }
    }
    void add(const void *p, size_t len) {
        const unsigned char *b = static_cast<const unsigned char *>(p);
        for (size_t k = 0; k < len; ++k) c = table[(c ^ b[k]) & 0xFFu] ^ (c >> 8);
    }
    uint64_t value() const { return (uint64_t)(c ^ 0xFFFFFFFFu); }
};

struct Writer {
    FILE *fp;
    Crc32 sum;
    bool ok;
    explicit Writer(FILE *f) : fp(f), ok(true) {}
    void put(const void *p, size_t len) {
        if (ok && len && std::fwrite(p, 1, len, fp) != len) ok = false;
        sum.add(p, len);
    }
    void i64(int64_t x) { put(&x, 8); }
    void u64(uint64_t x) { put(&x, 8); }
    void f64(double x) { put(&x, 8); }
    void vec(const std::vector<double> &a) { put(a.data(), a.size() * sizeof(double)); }
};

struct Reader {
    FILE *fp;
    Crc32 sum;
    bool ok;
    explicit Reader(FILE *f) : fp(f), ok(true) {}
    void get(void *p, size_t len) {
        if (ok && len && std::fread(p, 1, len, fp) != len) ok = false;
        if (ok) sum.add(p, len);
    }
    int64_t i64() { int64_t x = 0; get(&x, 8); return x; }
    uint64_t u64() { uint64_t x = 0; get(&x, 8); return x; }
    double f64() { double x = 0; get(&x, 8); return x; }
    void vec(std::vector<double> &a) { get(a.data(), a.size() * sizeof(double)); }
};

bool little_endian() {
    const uint16_t one = 1;
    unsigned char b;
    std::memcpy(&b, &one, 1);
    return b == 1;
}

}  // namespace

void H4Checkpoint::resize(uint64_t count) {
    n = count;
    r.assign(count * 4, 0.0);
    v.assign(count * 4, 0.0);
    f.assign(count * 6, 0.0);
    t.assign(count, 0.0);
    dt.assign(count, 0.0);
    a2.assign(count * 3, 0.0);
    a3.assign(count * 3, 0.0);
}

bool h4_write_checkpoint(const std::string &path, const H4Checkpoint &c, std::string *err) {
    if (!little_endian()) {
        if (err) *err = "checkpoints are little endian";
        return false;
    }
    const uint64_t n = c.n;
    if (c.r.size() != n * 4 || c.v.size() != n * 4 || c.f.size() != n * 6 || c.t.size() != n ||
        c.dt.size() != n || c.a2.size() != n * 3 || c.a3.size() != n * 3) {
        if (err) *err = "checkpoint arrays do not match n";
        return false;
    }
    const std::string tmp = path + ".tmp";
    FILE *fp = std::fopen(tmp.c_str(), "wb");
    if (!fp) {
        if (err) *err = "cannot open " + tmp + " for writing";
        return false;
    }
    Writer w(fp);
    w.put(MAGIC, 8);
    w.u64(n);
    w.f64(c.e2);
    w.f64(c.eta);
    w.f64(c.dt_min);
    w.f64(c.dt_max);
    w.i64(c.precision);
    w.f64(c.itime);
    w.f64(c.atime);
    w.f64(c.interval);
    w.i64(c.output_factor);
    w.i64(c.snapshot_number);
    w.f64(c.energy_ini);
    w.i64(c.iterations);
    w.i64(c.nsteps);
    w.i64(c.interactions);
    w.vec(c.r);
    w.vec(c.v);
    w.vec(c.f);
    w.vec(c.t);
    w.vec(c.dt);
    w.vec(c.a2);
    w.vec(c.a3);
    uint64_t h = w.sum.value();
    if (w.ok && std::fwrite(&h, 1, 8, fp) != 8) w.ok = false;
    if (std::fclose(fp) != 0) w.ok = false;
    if (!w.ok) {
        std::remove(tmp.c_str());
        if (err) *err = "write to " + tmp + " failed";
        return false;
    }
    if (std::rename(tmp.c_str(), path.c_str()) != 0) {
        std::remove(tmp.c_str());
        if (err) *err = "cannot rename " + tmp + " to " + path;
        return false;
    }
    return true;
}

bool h4_read_checkpoint(const std::string &path, H4Checkpoint &c, std::string *err) {
    if (!little_endian()) {
        if (err) *err = "checkpoints are little endian";
        return false;
    }
    FILE *fp = std::fopen(path.c_str(), "rb");
    if (!fp) {
        if (err) *err = "cannot open " + path;
        return false;
    }
    Reader rd(fp);
    char magic[8] = {0};
    rd.get(magic, 8);
    if (!rd.ok || std::memcmp(magic, MAGIC, 7) != 0) {
        std::fclose(fp);
        if (err) *err = path + " is not a Hermite4HIP checkpoint";
        return false;
    }
    if (magic[7] != MAGIC[7]) {
        std::fclose(fp);
        if (err) *err = path + ": unsupported checkpoint version";
        return false;
    }
    const uint64_t n = rd.u64();
    // 128 header bytes + 22 doubles per particle + checksum
    std::fseek(fp, 0, SEEK_END);
    const long size = std::ftell(fp);
    std::fseek(fp, 16, SEEK_SET);
    if (!rd.ok || n == 0 || n > (1ull << 32) || size < 0 ||
        (uint64_t)size != 128 + n * 22 * 8 + 8) {
        std::fclose(fp);
        if (err) *err = path + ": size does not match its particle count (truncated?)";
        return false;
    }
    c.resize(n);
    c.e2 = rd.f64();
    c.eta = rd.f64();
    c.dt_min = rd.f64();
    c.dt_max = rd.f64();
    c.precision = rd.i64();
    c.itime = rd.f64();
    c.atime = rd.f64();
    c.interval = rd.f64();
    c.output_factor = rd.i64();
    c.snapshot_number = rd.i64();
    c.energy_ini = rd.f64();
    c.iterations = rd.i64();
    c.nsteps = rd.i64();
    c.interactions = rd.i64();
    rd.vec(c.r);
    rd.vec(c.v);
    rd.vec(c.f);
    rd.vec(c.t);
    rd.vec(c.dt);
    rd.vec(c.a2);
    rd.vec(c.a3);
    uint64_t want = rd.sum.value(), got = 0;
    bool ok = rd.ok && std::fread(&got, 1, 8, fp) == 8;
    std::fclose(fp);
    if (!ok || got != want) {
        if (err) *err = path + ": checksum mismatch (corrupted checkpoint)";
        return false;
    }
    return true;
}
