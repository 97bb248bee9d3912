// compgt.cpp -- CompGT's command line (CompGT.cpp:134-163) over nbpg_compgt.
//
//   CompGT --base_path B.u8bin|B.i8bin --query_path Q --gt_path OUT.gt [--k 100]
//          [--engine mfma|plain] [--tie_rule reference|exact] [--use_mips]
//
// The element type follows the base file's extension as in the reference.
// --tie_rule reference (the default) keeps the reference's slot-replacement
// rule, so OUT.gt has the bytes the reference's CompGT writes; exact gives the
// (dist, id) top-k.  --use_mips ranks by the negated inner product
// (CompGT.cpp:51).  Float rows (.fbin/.bin) are not on this path and are
// refused.
#include <chrono>
#include <iostream>

#include "tool_io.h"

using namespace nbpg_tools;

int main(int argc, char **argv)
{
    try {
        Flags f(argc, argv);
        f.check({"base_path", "query_path", "gt_path", "k", "use_mips", "engine", "tie_rule"});
        const std::string bp = f.str("base_path"), qp = f.str("query_path"), gp = f.str("gt_path");
        if (bp.empty() || qp.empty() || gp.empty())
            throw std::runtime_error("--base_path, --query_path and --gt_path are required");
        int dtype;
        if (ends_with(bp, ".u8bin"))
            dtype = NBPG_U8;
        else if (ends_with(bp, ".i8bin"))
            dtype = NBPG_I8;
        else
            throw std::runtime_error("base rows must be .u8bin or .i8bin (float rows are not on this path)");
        const uint32_t k = (uint32_t)f.num("k", 100);
        const std::string engine = f.str("engine", "mfma"), rule = f.str("tie_rule", "reference");
        nbpg_ok(nbpg_init(0), "nbpg_init");
        Rows base = read_rows(bp), q = read_rows(qp);
        if (base.dims != q.dims)
            throw std::runtime_error("base and query dimensions differ");
        int mode = engine == "plain" ? NBPG_GT_PLAIN : NBPG_GT_MFMA;
        if (engine != "plain" && engine != "mfma")
            throw std::runtime_error("--engine mfma|plain");
        // the MFMA tiles take 16-byte rows; other widths use the plain engine
        if (mode == NBPG_GT_MFMA && base.dims % 16)
            mode = NBPG_GT_PLAIN;
        if (rule == "reference")
            mode |= NBPG_GT_REFTIES;
        else if (rule != "exact")
            throw std::runtime_error("--tie_rule reference|exact");
        if (f.has("use_mips"))
            mode |= NBPG_GT_MIPS;
        DevBuf db(base.data.data(), base.data.size()), dq(q.data.data(), q.data.size());
        DevBuf di((size_t)q.n * k * 4), dd((size_t)q.n * k * 4);
        const auto t0 = std::chrono::steady_clock::now();
        nbpg_ok(nbpg_compgt(dtype, db.p, base.n, dq.p, q.n, base.dims, k, mode, (uint32_t *)di.p,
                            (float *)dd.p, nullptr),
                "nbpg_compgt");
        hip_ok(hipDeviceSynchronize(), "compgt");
        const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::vector<uint32_t> ids((size_t)q.n * k);
        std::vector<float> d((size_t)q.n * k);
        di.download(ids.data(), ids.size() * 4);
        dd.download(d.data(), d.size() * 4);
        write_gt(gp, q.n, k, ids, d);
        std::cout << "ground truth of " << q.n << " queries over " << base.n << " rows, k = " << k
                  << " in " << s << " s" << std::endl;
        return 0;
    } catch (const std::exception &e) {
        std::cerr << "CompGT: " << e.what() << std::endl;
        return 1;
    }
}
