// pcacompress.cpp -- PCACompress's command line (PCACompress.cpp:21-110) over
// nbpg_pca_project.
//
//   PCACompress --ifile X.fbin|X.u8bin|X.i8bin --ofile Y.{fbin,u8bin,i8bin,u4bin,i4bin}
//               --matFile W.mat [--index_type uint32] [--mode plain|bf16]
//
// The output extension picks the code type and chunk size as in the reference
// (PCACompress.cpp:25-37); rows stream through the device in blocks of
// 100000 (PCACompress.cpp:55-76).  --mode plain (the default) sums in the
// reference's order, so Y has the reference's bytes; bf16 uses the matrix
// cores.
#include <chrono>
#include <iostream>

#include "tool_io.h"

using namespace nbpg_tools;

int main(int argc, char **argv)
{
    try {
        Flags f(argc, argv);
        f.check({"ifile", "ofile", "matFile", "index_type", "mode"});
        const std::string in = f.str("ifile"), out = f.str("ofile"), mp = f.str("matFile");
        if (in.empty() || out.empty() || mp.empty())
            throw std::runtime_error("--ifile, --ofile and --matFile are required");
        if (f.str("index_type", "uint32") != "uint32")
            throw std::runtime_error("--index_type: only uint32 headers");
        int xt;
        if (ends_with(in, ".fbin") || ends_with(in, ".bin"))
            xt = NBPG_F32;
        else if (ends_with(in, ".u8bin"))
            xt = NBPG_U8;
        else if (ends_with(in, ".i8bin"))
            xt = NBPG_I8;
        else
            throw std::runtime_error("input must be .fbin, .bin, .u8bin or .i8bin");
        int code;
        uint32_t chunk = 1;
        if (ends_with(out, ".u4bin"))
            code = NBPG_CODE_U4, chunk = 2;
        else if (ends_with(out, ".i4bin"))
            code = NBPG_CODE_I4, chunk = 2;
        else if (ends_with(out, ".u8bin"))
            code = NBPG_CODE_U8;
        else if (ends_with(out, ".i8bin"))
            code = NBPG_CODE_I8;
        else
            code = NBPG_CODE_F32; // PCACompress.cpp:32: anything else is written as floats
        const std::string m = f.str("mode", "plain");
        if (m != "plain" && m != "bf16")
            throw std::runtime_error("--mode plain|bf16");
        const int mode = m == "plain" ? NBPG_PCA_PLAIN : NBPG_PCA_BF16;
        uint32_t d_out = 0, d_in = 0;
        std::vector<float> W, bias;
        read_mat(mp, d_out, d_in, W, bias);
        nbpg_ok(nbpg_init(0), "nbpg_init");
        Rows x = read_rows(in);
        if (x.dims != d_in)
            throw std::runtime_error("input dimension differs from the matrix's");
        const uint32_t out_dims = d_out / chunk;
        std::ofstream w(out, std::ios::binary);
        if (!w)
            throw std::runtime_error("cannot write " + out);
        w.write(reinterpret_cast<const char *>(&x.n), 4);
        w.write(reinterpret_cast<const char *>(&out_dims), 4);
        DevBuf dW(W.data(), W.size() * 4), db(bias.data(), bias.size() * 4);
        const uint64_t B = 100000;
        DevBuf dx(std::min<uint64_t>(B, x.n) * x.row_bytes());
        DevBuf dy(code == NBPG_CODE_F32 ? std::min<uint64_t>(B, x.n) * d_out * 4 : 0);
        DevBuf dc(code == NBPG_CODE_F32 ? 0 : std::min<uint64_t>(B, x.n) * out_dims);
        std::vector<uint8_t> host(std::min<uint64_t>(B, x.n) * std::max<uint64_t>(d_out * 4ull, out_dims));
        const auto t0 = std::chrono::steady_clock::now();
        for (uint64_t s = 0; s < x.n; s += B) {
            const uint64_t m_rows = std::min<uint64_t>(B, x.n - s);
            hip_ok(hipMemcpy(dx.p, x.data.data() + s * x.row_bytes(), m_rows * x.row_bytes(),
                             hipMemcpyHostToDevice),
                   "upload");
            const bool fl = code == NBPG_CODE_F32;
            nbpg_ok(nbpg_pca_project(xt, dx.p, m_rows, d_in, (const float *)dW.p, (const float *)db.p,
                                     d_out, mode, code, fl ? nullptr : (uint8_t *)dc.p,
                                     fl ? (float *)dy.p : nullptr, nullptr),
                    "nbpg_pca_project");
            const size_t ob = code == NBPG_CODE_F32 ? m_rows * d_out * 4 : m_rows * out_dims;
            (code == NBPG_CODE_F32 ? dy : dc).download(host.data(), ob);
            w.write(reinterpret_cast<const char *>(host.data()), (std::streamsize)ob);
        }
        const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::cout << "[Timer 1]:compress pca:" << sec << std::endl;
        return 0;
    } catch (const std::exception &e) {
        std::cerr << "PCACompress: " << e.what() << std::endl;
        return 1;
    }
}
