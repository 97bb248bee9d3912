// vamana.cpp -- the reference's Vamana command line (Vamana.cpp:14-340) over the
// C-ABI: build (or load) the graph on the device, save <index_path>.graph,
// then, given queries and ground truth, the recall / QPS sweep of
// check_nn_recall.h:194-268 with its CSV (check_nn_recall.h:12-48).
//
//   Vamana --quant_base_path B.{u8bin,i8bin,u4bin,i4bin} --index_path P [--rebuild]
//          [--pre_index_path P0] [--R 50] [--L 70] [--VL 140] [--limit 0] [--top_degree 0]
//          [--num_passes 2] [--num_split 4] [--dynamic_prune] [--rand_sp]
//          [--min_alpha 1] [--max_alpha 1.1] [--min_cos_angle 0] [--max_cos_angle 0]
//          [--quant_query_path Q --gt_path G.gt [--res_path R.csv] [--k 10] [--Q 0]
//           [--q_num 0] [--verbose]]
//          [--search_mode ref|fast] [--sp_seed 1]
//
//          [--base_path X.{fbin,bin,u8bin,i8bin,u4bin,i4bin} --query_path XQ
//           [--rerank_factor 100] [--cache_raw]]
//
// --search_mode ref (the default) returns the reference's IDs and counters;
// fast runs the persistent kernel.  With --base_path / --query_path (raw
// rows of another size than the quantized ones) the sweep searches the
// quantized rows and takes final distances from the raw ones
// (beam_search_rerank, beamSearch.h:183-247): --rerank_factor > 0 re-scores
// the frontier head, 0 is the filtering mode (re-scores the visited list;
// with --search_mode fast for beams up to 256); the raw rows stay in device
// memory (--cache_raw is accepted and changes nothing).
// 4-split builds draw their random start points from --sp_seed instead of
// std::random_device.  --use_mips builds and searches with the negated
// inner product (DistanceMetric::MIPS).  Not on this path (refused):
// --dir_bias_scale > 0.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <iomanip>
#include <iostream>
#include <set>
#include <sstream>
#include <sys/stat.h>

#include "tool_io.h"

using namespace nbpg_tools;

namespace {

bool file_exists(const std::string &p)
{
    struct stat st;
    return stat(p.c_str(), &st) == 0;
}

// point.h:35-156 on the host, for the tie extension of the recall: squared
// L2, or with mips the negated inner product
float host_dist(int dtype, const uint8_t *a, const uint8_t *b, uint32_t dims, bool mips)
{
    if (dtype == NBPG_F32) { // point.h:137-156, summed in index order
        const float *p = reinterpret_cast<const float *>(a), *q = reinterpret_cast<const float *>(b);
        float r = 0.0f;
        for (uint32_t i = 0; i < dims; ++i) {
            if (mips) {
                r += p[i] * q[i];
            } else {
                const float d = p[i] - q[i];
                r += d * d;
            }
        }
        return mips ? -r : r;
    }
    int32_t s = 0;
    for (uint32_t i = 0; i < dims; ++i) {
        int xa, xb, ya = 0, yb = 0; // the element pair, or the two nibble pairs
        if (dtype == NBPG_U8) {
            xa = a[i];
            xb = b[i];
        } else if (dtype == NBPG_I8) {
            xa = (int8_t)a[i];
            xb = (int8_t)b[i];
        } else {
            xa = a[i] & 0xF0;
            xb = b[i] & 0xF0;
            ya = (a[i] << 4) & 0xFF;
            yb = (b[i] << 4) & 0xFF;
        }
        s += mips ? xa * xb + ya * yb : (xa - xb) * (xa - xb) + (ya - yb) * (ya - yb);
    }
    return mips ? -(float)s : (float)s;
}

struct Result {
    double recall;
    uint32_t avg_cmps, tail_cmps, avg_visited, tail_visited;
    float qps;
    uint32_t k, Q;
    float cut;
    uint64_t limit;
    uint32_t degree_limit;
    uint64_t nq;
    void print() const
    {
        std::cout << "For " << k << "@" << k << " recall = " << recall << ", QPS = " << qps
                  << ", Q = " << Q << ", cut = " << cut << ", visited limit = " << limit
                  << ", degree limit: " << degree_limit << ", average visited = " << avg_visited
                  << ", average cmps = " << avg_cmps << std::endl;
    }
};

// query_params.h:66-76
std::pair<uint32_t, uint32_t> statistics(std::vector<uint32_t> s)
{
    std::sort(s.begin(), s.end());
    double sum = 0;
    for (auto v : s)
        sum += v;
    const uint32_t avg = (uint32_t)(sum / s.size());
    const size_t tail = (size_t)(.99 * (float)s.size());
    return {avg, s[tail]};
}

std::string csv_str(const std::string &v)
{
    std::string o = "\"";
    for (char c : v) {
        if (c == '"')
            o += '"';
        o += c;
    }
    return o + "\"";
}

} // namespace

int main(int argc, char **argv)
{
    try {
        Flags f(argc, argv);
        f.check({"quant_base_path", "pre_index_path", "index_path", "R", "L", "limit", "VL",
                 "top_degree", "num_passes", "num_split", "dir_bias_scale", "dynamic_prune",
                 "use_mips", "rebuild", "rand_sp", "max_alpha", "min_alpha", "max_cos_angle",
                 "min_cos_angle", "cache_raw", "base_path", "query_path", "quant_query_path",
                 "gt_path", "res_path", "k", "verbose", "Q", "rerank_factor", "q_num",
                 "search_mode", "sp_seed"});
        const std::string bp = f.str("quant_base_path"), ip = f.str("index_path");
        if (bp.empty() || ip.empty())
            throw std::runtime_error("--quant_base_path and --index_path are required");
        const bool mips = f.has("use_mips");
        if (f.num("dir_bias_scale", 0) > 0)
            throw std::runtime_error("--dir_bias_scale > 0 is not on this path");
        int dtype;
        if (ends_with(bp, ".u8bin"))
            dtype = NBPG_U8;
        else if (ends_with(bp, ".i8bin"))
            dtype = NBPG_I8;
        else if (ends_with(bp, ".u4bin"))
            dtype = NBPG_U4X2;
        else if (ends_with(bp, ".i4bin"))
            dtype = NBPG_I4X2;
        else
            throw std::runtime_error("--quant_base_path must be .u8bin, .i8bin, .u4bin or .i4bin");
        nbpg_ok(nbpg_init(0), "nbpg_init");
        Rows base = read_rows(bp);
        const uint32_t R = (uint32_t)f.num("R", 50);
        nbpg_build_params p;
        p.R = R;
        p.L = (uint32_t)f.num("L", 70);
        p.VL = (uint32_t)f.num("VL", 140);
        p.top_degree = (uint32_t)f.num("top_degree", 0);
        p.num_split = (uint32_t)f.num("num_split", 4);
        p.num_passes = (uint32_t)f.num("num_passes", 2);
        p.limit = (uint64_t)f.num("limit", 0);
        p.min_alpha = f.num("min_alpha", 1.0);
        p.max_alpha = f.num("max_alpha", 1.1);
        p.min_cos_angle = f.num("min_cos_angle", 0.0);
        p.max_cos_angle = f.num("max_cos_angle", 0.0);
        p.dynamic_prune = f.has("dynamic_prune");
        p.rand_sp = f.has("rand_sp");
        p.repair_isolate = 1;
        p.sp_seed = (uint64_t)f.num("sp_seed", 1);

        const std::string gfile = ip + ".graph";
        nbpg_index *ix = nullptr;
        uint64_t n = base.n;
        uint32_t gR = R;
        double build_s = 0;
        uint64_t isolated = 0;
        // the graph moves between the file and the device without a host
        // copy of its rows (nbpg_index_load_graph / nbpg_index_save_graph)
        if (file_exists(gfile) && !f.has("rebuild")) {
            read_graph_degrees(gfile, n, gR);
            if (n != base.n)
                throw std::runtime_error("graph and base sizes differ");
            nbpg_ok(nbpg_index_create(&ix, dtype, n, base.dims, gR), "nbpg_index_create");
            nbpg_ok(nbpg_index_set_metric(ix, mips ? NBPG_METRIC_MIPS : NBPG_METRIC_L2), "set_metric");
            nbpg_ok(nbpg_index_set_points(ix, base.data.data(), base.row_bytes(), 0, n), "set_points");
            nbpg_ok(nbpg_index_load_graph(ix, gfile.c_str()), "load_graph");
        } else {
            nbpg_ok(nbpg_index_create(&ix, dtype, n, base.dims, R), "nbpg_index_create");
            nbpg_ok(nbpg_index_set_metric(ix, mips ? NBPG_METRIC_MIPS : NBPG_METRIC_L2), "set_metric");
            nbpg_ok(nbpg_index_set_points(ix, base.data.data(), base.row_bytes(), 0, n), "set_points");
            if (!f.str("pre_index_path").empty()) { // Vamana.cpp:174-177
                const std::string pf = f.str("pre_index_path") + ".graph";
                uint64_t pn = 0;
                uint32_t pR = 0;
                read_graph_degrees(pf, pn, pR);
                if (pn != n || pR != R)
                    throw std::runtime_error("pre-built graph shape differs");
                nbpg_ok(nbpg_index_load_graph(ix, pf.c_str()), "load_graph");
            }
            const auto t0 = std::chrono::steady_clock::now();
            nbpg_ok(nbpg_build(ix, &p, &isolated), "nbpg_build");
            build_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            std::cout << "Writing graph with " << n << " points and max degree " << R << std::endl;
            nbpg_ok(nbpg_index_save_graph(ix, gfile.c_str()), "save_graph");
        }
        // graph statistics (parse_results.h:7-23, Vamana.cpp:91-111)
        uint64_t sum = 0;
        uint32_t maxd = 0;
        {
            uint64_t fn = 0;
            uint32_t fR = 0;
            for (uint32_t d : read_graph_degrees(gfile, fn, fR)) {
                sum += d;
                maxd = std::max(maxd, d);
            }
        }
        const double avg_deg = (double)sum / (double)n;
        std::ostringstream params;
        params << "R = " << gR << ", L = " << p.L << ", alpha = (" << std::fixed << std::setprecision(2)
               << p.min_alpha << "," << p.max_alpha << "), use_mips = " << (mips ? 1 : 0);
        std::cout << "Vamana graph built with " << n << " points and parameters " << params.str()
                  << std::endl;
        std::cout << "Graph has average degree " << avg_deg << ", popular average degree 0"
                  << " and maximum degree " << maxd << std::endl;
        std::cout << "Isolated nodes num is " << isolated << std::endl;
        std::cout << "Graph built in " << build_s << " seconds" << std::endl;

        const std::string qp = f.str("quant_query_path"), gp = f.str("gt_path");
        if (qp.empty() || gp.empty()) {
            nbpg_index_destroy(ix);
            return 0;
        }
        Rows q = read_rows(qp);
        const uint64_t qn = f.num("q_num", 0) > 0 ? std::min<uint64_t>(q.n, (uint64_t)f.num("q_num", 0)) : q.n;
        uint32_t gnq = 0, gk = 0;
        std::vector<uint32_t> gt;
        read_gt(gp, gnq, gk, gt);
        if (gnq < qn)
            throw std::runtime_error("ground truth has fewer queries");
        const uint32_t kk = f.num("k", 0) > 0 ? (uint32_t)f.num("k", 0) : 10;
        if (kk > gk)
            throw std::runtime_error("k too large for the ground truth");
        const bool verbose = f.has("verbose");
        const std::string sm = f.str("search_mode", "ref");
        if (sm != "ref" && sm != "fast")
            throw std::runtime_error("--search_mode ref|fast");
        const int mode = sm == "ref" ? NBPG_SEARCH_REF : NBPG_SEARCH_FAST;
        DevBuf dq(q.data.data(), qn * q.row_bytes());

        // raw rows (Vamana.cpp:212-272): recall always measures against them;
        // the search uses them when their row size differs from the
        // quantized rows' (beamSearch.h:21-38)
        const std::string xbp = f.str("base_path"), xqp = f.str("query_path");
        if (xbp.empty() != xqp.empty())
            throw std::runtime_error("--base_path and --query_path go together");
        Rows xb, xq;
        int xdtype = dtype;
        if (!xbp.empty()) {
            if (ends_with(xbp, ".u8bin"))
                xdtype = NBPG_U8;
            else if (ends_with(xbp, ".i8bin"))
                xdtype = NBPG_I8;
            else if (ends_with(xbp, ".u4bin"))
                xdtype = NBPG_U4X2;
            else if (ends_with(xbp, ".i4bin"))
                xdtype = NBPG_I4X2;
            else if (ends_with(xbp, ".fbin") || ends_with(xbp, ".bin"))
                xdtype = NBPG_F32;
            else
                throw std::runtime_error("--base_path must be .u8bin, .i8bin, .u4bin, .i4bin, .fbin or .bin");
            xb = read_rows(xbp);
            xq = read_rows(xqp, qn);
            if (xb.n != n || xq.n < qn || xq.dims != xb.dims || xq.elem != xb.elem)
                throw std::runtime_error("raw rows do not match the quantized rows");
        }
        const Rows &rb = xbp.empty() ? base : xb;
        const Rows &rq = xbp.empty() ? q : xq;
        const bool exact = !xbp.empty() && xb.row_bytes() != base.row_bytes();
        DevBuf dxb(exact ? xb.data.size() : 0), dxq(exact ? qn * xq.row_bytes() : 0);
        if (exact) {
            hip_ok(hipMemcpy(dxb.p, xb.data.data(), dxb.bytes, hipMemcpyHostToDevice), "upload");
            hip_ok(hipMemcpy(dxq.p, xq.data.data(), dxq.bytes, hipMemcpyHostToDevice), "upload");
        }
        const uint32_t flag_rf = (uint32_t)f.num("rerank_factor", 100);

        // checkRecall (check_nn_recall.h:51-80): search with qk results,
        // recall over the first k
        auto check = [&](uint32_t k, uint32_t qk, uint32_t Q, double cut, uint64_t limit,
                         uint32_t dlim, uint32_t rf) -> Result {
            DevBuf di(qn * qk * 4), dd(qn * qk * 4), ds(qn * 8);
            nbpg_query_params qpm = {qk, cut, Q, limit, dlim};
            const uint32_t sp = 0;
            const auto t0 = std::chrono::steady_clock::now();
            if (exact)
                nbpg_ok(nbpg_search_rerank(ix, dq.p, qn, q.row_bytes(), xdtype, dxb.p, xb.dims,
                                           dxq.p, xq.row_bytes(), &sp, 1, &qpm, rf, mode, nullptr,
                                           (uint32_t *)di.p, (float *)dd.p, (uint32_t *)ds.p,
                                           nullptr),
                        "nbpg_search_rerank");
            else
                nbpg_ok(nbpg_search(ix, dq.p, qn, q.row_bytes(), &sp, 1, &qpm, mode, nullptr,
                                    (uint32_t *)di.p, (float *)dd.p, (uint32_t *)ds.p, nullptr),
                        "nbpg_search");
            hip_ok(hipDeviceSynchronize(), "search");
            const double t = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            std::vector<uint32_t> ids(qn * qk), st(qn * 2);
            di.download(ids.data(), ids.size() * 4);
            ds.download(st.data(), st.size() * 4);
            // recall with ties (check_nn_recall.h:126-167)
            uint64_t correct = 0;
            for (uint64_t i = 0; i < qn; ++i) {
                const uint8_t *qrow = rq.data.data() + i * rq.row_bytes();
                auto dist = [&](uint32_t id) {
                    return host_dist(xdtype, rb.data.data() + (uint64_t)id * rb.row_bytes(), qrow,
                                     rb.dims, mips);
                };
                std::vector<uint32_t> with_ties(gt.begin() + i * gk, gt.begin() + i * gk + k);
                const float last = dist(gt[i * gk + k - 1]);
                for (uint32_t l = k; l < gk; ++l) {
                    const uint32_t g = gt[i * gk + l];
                    if (dist(g) == last)
                        with_ties.push_back(g);
                    else
                        break;
                }
                std::set<uint32_t> rep;
                for (uint32_t l = 0; l < k && l < qk; ++l)
                    if (ids[i * qk + l] != 0xFFFFFFFFu)
                        rep.insert(ids[i * qk + l]);
                for (uint32_t g : with_ties)
                    correct += rep.count(g);
            }
            std::vector<uint32_t> vis(qn), cmp(qn);
            for (uint64_t i = 0; i < qn; ++i) {
                vis[i] = st[2 * i];
                cmp[i] = st[2 * i + 1];
            }
            auto dsx = statistics(cmp), vsx = statistics(vis);
            // a float ratio, as check_nn_recall.h:165 computes it
            const float recall = static_cast<float>(correct) / static_cast<float>(k * qn);
            Result r{recall, dsx.first, dsx.second, vsx.first, vsx.second,
                     (float)(qn / t), k, Q, (float)cut, limit, dlim, qn};
            if (verbose)
                std::cout << "search: Q=" << Q << ", k=" << qk << ", limit=" << limit << ", recall="
                          << recall << ", visited=" << r.avg_visited << ", comparisons="
                          << r.avg_cmps << ", QPS=" << r.qps << std::endl;
            return r;
        };

        const uint32_t fixedQ = (uint32_t)f.num("Q", 0);
        if (fixedQ) {
            for (int i = 0; i < 5; ++i)
                check(kk, kk, fixedQ, 1.35, n, gR, flag_rf).print();
            nbpg_index_destroy(ix);
            return 0;
        }
        // check_nn_recall.h:206-257
        std::vector<Result> results;
        const uint32_t beams[] = {10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 22, 24, 26, 28,
                                  30, 32, 34, 36, 38, 40, 45, 50, 55, 60, 65, 70, 80, 90, 100,
                                  120, 140, 160, 180, 200, 225, 250, 275, 300, 375, 500, 750, 1000};
        for (uint32_t Q : beams)
            if (Q >= kk)
                results.push_back(check(kk, kk, Q, 1.35, n, gR, flag_rf));
        // the limited and best-accuracy points rebuild QueryParams with its
        // default rerank factor 100 (check_nn_recall.h:236-248, query_params.h:34-35)
        for (uint32_t l : {10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 28, 30, 35})
            results.push_back(check(kk, kk, std::max(l, kk), 1.35, l, std::min(gR, 5 * l), 100));
        results.push_back(check(kk, 100, 1000, 10.0, n, gR, 100));
        const float buckets[] = {.1f, .2f, .3f, .4f, .5f, .6f, .7f, .75f, .8f, .85f, .9f, .93f, .95f,
                                 .97f, .98f, .99f, .995f, .999f, .9995f, .9999f, .99995f, .99999f};
        const int nb = sizeof(buckets) / sizeof(buckets[0]);
        std::vector<Result> picked;
        std::vector<float> pb;
        for (int i = 0; i < nb; ++i) { // parse_results.h:140-178
            std::vector<Result> cand;
            for (auto &r : results)
                if (r.recall >= buckets[i] && (i == nb - 1 || r.recall <= buckets[i + 1]))
                    cand.push_back(r);
            if (cand.empty())
                continue;
            auto best = *std::max_element(cand.begin(), cand.end(),
                                          [](const Result &a, const Result &b) { return a.qps < b.qps; });
            best.print();
            picked.push_back(best);
            pb.push_back(buckets[i]);
        }
        std::cout << std::endl;
        const std::string rp = f.str("res_path");
        if (!rp.empty()) { // check_nn_recall.h:12-48, appended
            std::ofstream csv(rp, std::ios::app);
            csv << csv_str("GRAPH") << "," << csv_str("Parameters") << "," << csv_str("Size") << ","
                << csv_str("Build time") << "," << csv_str("Avg degree") << "," << csv_str("Max degree")
                << "," << csv_str("Isolated") << "," << csv_str("Popular Avg degree") << std::endl;
            csv << csv_str("Vamana") << "," << csv_str(params.str()) << "," << n << "," << build_s << ","
                << avg_deg << "," << maxd << "," << isolated << "," << 0 << std::endl;
            csv << std::endl;
            csv << csv_str("Num queries") << "," << csv_str("Target recall") << ","
                << csv_str("Actual recall") << "," << csv_str("QPS") << "," << csv_str("Average Cmps")
                << "," << csv_str("Tail Cmps") << "," << csv_str("Average Visited") << ","
                << csv_str("Tail Visited") << "," << csv_str("k") << "," << csv_str("Q") << ","
                << csv_str("cut") << std::endl;
            for (size_t i = 0; i < picked.size(); ++i) {
                const Result &r = picked[i];
                csv << r.nq << "," << pb[i] << "," << r.recall << "," << r.qps << "," << r.avg_cmps << ","
                    << r.tail_cmps << "," << r.avg_visited << "," << r.tail_visited << "," << r.k << ","
                    << r.Q << "," << r.cut << std::endl;
            }
            csv << std::endl << std::endl;
        }
        nbpg_index_destroy(ix);
        return 0;
    } catch (const std::exception &e) {
        std::cerr << "Vamana: " << e.what() << std::endl;
        return 1;
    }
}
