// tool_io.h -- file formats and flag parsing shared by the command-line tools
// (Vamana, CompGT, PCACompress) that drive the C-ABI from the host.
//
// Formats, as the reference reads and writes them:
//   .u8bin/.i8bin/.fbin/.u4bin/.i4bin  [u32 n][u32 dims][n*dims elements]
//                                      (point_range.h:17-30; 4-bit: dims = bytes)
//   .graph  [u32 n][u32 R][u32 deg[n]][neighbours in node order] (graph.h:168-221, 339-379)
//   .gt     [u32 nq][u32 k][u32 ids nq*k][f32 dists nq*k] (CompGT.cpp:126-133)
//   .mat    [u32 d_out][u32 d_in][u8 has_bias][f32 W][f32 bias] (matrix.h:59-83)
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "nbpg.h"

namespace nbpg_tools {

inline bool ends_with(const std::string &s, const std::string &t)
{
    return s.size() >= t.size() && s.compare(s.size() - t.size(), t.size(), t) == 0;
}

// --name value and --flag (a flag is followed by another --option or nothing)
struct Flags {
    std::map<std::string, std::string> kv;
    Flags(int argc, char **argv)
    {
        for (int i = 1; i < argc; ++i) {
            std::string a = argv[i];
            if (a.rfind("--", 0) != 0)
                throw std::runtime_error("unexpected argument " + a);
            const std::string key = a.substr(2);
            if (i + 1 < argc && std::string(argv[i + 1]).rfind("--", 0) != 0)
                kv[key] = argv[++i];
            else
                kv[key] = "1";
        }
    }
    bool has(const std::string &k) const { return kv.count(k) != 0; }
    std::string str(const std::string &k, const std::string &d = "") const
    {
        auto it = kv.find(k);
        return it == kv.end() ? d : it->second;
    }
    double num(const std::string &k, double d) const
    {
        auto it = kv.find(k);
        return it == kv.end() ? d : std::stod(it->second);
    }
    // flags the tool does not know are errors, as CLI11 makes them
    void check(const std::vector<std::string> &known) const
    {
        for (auto &p : kv) {
            bool ok = false;
            for (auto &k : known)
                ok |= p.first == k;
            if (!ok)
                throw std::runtime_error("unknown option --" + p.first);
        }
    }
};

struct Rows {
    uint32_t n = 0, dims = 0; // dims as stored in the header
    uint32_t elem = 1;        // bytes per element
    std::vector<uint8_t> data;
    uint64_t row_bytes() const { return (uint64_t)dims * elem; }
};

inline uint32_t elem_size_of(const std::string &path)
{
    return (ends_with(path, ".fbin") || ends_with(path, ".bin")) ? 4 : 1;
}

inline Rows read_rows(const std::string &path, uint64_t max_rows = 0)
{
    std::ifstream f(path, std::ios::binary);
    if (!f)
        throw std::runtime_error("cannot open " + path);
    Rows r;
    f.read(reinterpret_cast<char *>(&r.n), 4);
    f.read(reinterpret_cast<char *>(&r.dims), 4);
    r.elem = elem_size_of(path);
    if (max_rows && r.n > max_rows)
        r.n = (uint32_t)max_rows;
    r.data.resize((size_t)r.n * r.row_bytes());
    f.read(reinterpret_cast<char *>(r.data.data()), (std::streamsize)r.data.size());
    if (!f)
        throw std::runtime_error("short file " + path);
    return r;
}

// the header and degree array of a .graph file (graph.h:182-190)
inline std::vector<uint32_t> read_graph_degrees(const std::string &path, uint64_t &n, uint32_t &R)
{
    std::ifstream f(path, std::ios::binary);
    if (!f)
        throw std::runtime_error("cannot open " + path);
    uint32_t hn = 0, hr = 0;
    f.read(reinterpret_cast<char *>(&hn), 4);
    f.read(reinterpret_cast<char *>(&hr), 4);
    n = hn;
    R = hr;
    std::vector<uint32_t> deg(n);
    f.read(reinterpret_cast<char *>(deg.data()), (std::streamsize)(n * 4));
    if (!f)
        throw std::runtime_error("short file " + path);
    return deg;
}


inline void write_gt(const std::string &path, uint32_t nq, uint32_t k, const std::vector<uint32_t> &ids,
                     const std::vector<float> &d)
{
    std::ofstream f(path, std::ios::binary);
    if (!f)
        throw std::runtime_error("cannot write " + path);
    f.write(reinterpret_cast<const char *>(&nq), 4);
    f.write(reinterpret_cast<const char *>(&k), 4);
    f.write(reinterpret_cast<const char *>(ids.data()), (std::streamsize)ids.size() * 4);
    f.write(reinterpret_cast<const char *>(d.data()), (std::streamsize)d.size() * 4);
}

inline void read_gt(const std::string &path, uint32_t &nq, uint32_t &k, std::vector<uint32_t> &ids)
{
    std::ifstream f(path, std::ios::binary);
    if (!f)
        throw std::runtime_error("cannot open " + path);
    f.read(reinterpret_cast<char *>(&nq), 4);
    f.read(reinterpret_cast<char *>(&k), 4);
    ids.resize((size_t)nq * k);
    f.read(reinterpret_cast<char *>(ids.data()), (std::streamsize)ids.size() * 4);
    if (!f)
        throw std::runtime_error("short file " + path);
}

inline void read_mat(const std::string &path, uint32_t &d_out, uint32_t &d_in, std::vector<float> &W,
                     std::vector<float> &bias)
{
    std::ifstream f(path, std::ios::binary);
    if (!f)
        throw std::runtime_error("cannot open " + path);
    uint8_t has_bias = 0;
    f.read(reinterpret_cast<char *>(&d_out), 4);
    f.read(reinterpret_cast<char *>(&d_in), 4);
    f.read(reinterpret_cast<char *>(&has_bias), 1);
    W.resize((size_t)d_out * d_in);
    f.read(reinterpret_cast<char *>(W.data()), (std::streamsize)W.size() * 4);
    bias.assign(d_out, 0.0f);
    if (has_bias)
        f.read(reinterpret_cast<char *>(bias.data()), (std::streamsize)d_out * 4);
    if (!f)
        throw std::runtime_error("short file " + path);
}

inline void hip_ok(hipError_t e, const char *what)
{
    if (e != hipSuccess)
        throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

inline void nbpg_ok(int rc, const char *what)
{
    if (rc != NBPG_OK)
        throw std::runtime_error(std::string(what) + ": " + nbpg_last_error());
}

// device copy of a host buffer, freed with the object
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    explicit DevBuf(size_t b) : bytes(b) { hip_ok(hipMalloc(&p, b ? b : 1), "hipMalloc"); }
    DevBuf(const void *host, size_t b) : DevBuf(b)
    {
        hip_ok(hipMemcpy(p, host, b, hipMemcpyHostToDevice), "upload");
    }
    ~DevBuf()
    {
        if (p)
            hipFree(p);
    }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    void download(void *host, size_t b) const
    {
        hip_ok(hipMemcpy(host, p, b, hipMemcpyDeviceToHost), "download");
    }
};

} // namespace nbpg_tools
