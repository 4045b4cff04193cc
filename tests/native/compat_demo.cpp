// Drives include/bs_compat.hpp the way the reference's drivers use their
// entry points: the FastFlow map (run/wait, 100 passes), bs_thread over bound
// globals, the SSE pointer-batch form, the per-option forms, the ParallelFor
// front (100 parallel_for_thid passes) and the SkePU Vector/Map front (100
// map calls).  Writes one "%.9g" line per price of each form, then a "#"
// line with the SkePU vectors' transfer counts, for tests/test_tools.py.
//   compat_demo <input: N then N lines "s k r v t type"> <output>
#include <stdio.h>

#include <vector>

#include "bs_compat.hpp"

int main(int argc, char** argv) {
    if (argc != 3) return 2;
    FILE* f = fopen(argv[1], "r");
    if (!f) return 2;
    int n = 0;
    if (fscanf(f, "%d", &n) != 1 || n < 1) return 2;
    std::vector<float> s(n), k(n), r(n), v(n), t(n);
    std::vector<int> ty(n);
    for (int i = 0; i < n; ++i)
        if (fscanf(f, "%f %f %f %f %f %d", &s[i], &k[i], &r[i], &v[i], &t[i], &ty[i]) != 6) return 2;
    fclose(f);
    std::vector<float> p_map(n), p_thread(n, -1.0f), p_batch(n), p_scalar(n), cndf(n);
    std::vector<float> p_pf(n), p_sk(n);
    int up_s = -1, down_p = -1, refused = 0;
    try {
        {
            bs_compat::Map m(n, s.data(), k.data(), r.data(), v.data(), t.data(), ty.data(),
                             p_map.data());
            m.run();
            m.wait();
        }
        bs_compat::bind_globals(n, 2, s.data(), k.data(), r.data(), v.data(), t.data(), ty.data(),
                                p_thread.data());
        for (int tid = 0; tid < 2; ++tid) bs_compat::bs_thread(&tid);
        bs_compat::BlkSchlsEqEuroNoDiv(p_batch.data(), n, s.data(), k.data(), r.data(), v.data(),
                                       t.data(), ty.data(), 0.0f);
        for (int i = 0; i < n && i < 64; ++i) {
            p_scalar[i] = bs_compat::BlkSchlsEqEuroNoDiv(s[i], k[i], r[i], v[i], t[i], ty[i], 0.0f);
            cndf[i] = bs_compat::CNDF((float)(i - 32) / 8.0f);
        }
        {
            // FastFlow: the map's svc loop over a ParallelFor (blackscholes.c:316-347)
            bs_compat::PriceBody body(n, s.data(), k.data(), r.data(), v.data(), t.data(),
                                      ty.data(), p_pf.data());
            bs_compat::ParallelFor pf(2);
            for (int run = 0; run < bs_compat::kNumRuns; ++run)
                pf.parallel_for_thid(0, n, 1, 0, body, 2);
        }
        {
            // SkePU2 (blackscholes_skepu_omp.cpp:710-728)
            namespace sk = bs_compat::skepu;
            sk::Vector<float> sv(s.data(), n), kv(k.data(), n), rv(r.data(), n), vv(v.data(), n),
                tv(t.data(), n), pv(p_sk.data(), n);
            sk::Vector<int> tyv(ty.data(), n);
            sk::Vector<char> data(n);
            sk::Map map;
            try {
                map.setBackend(sk::BackendSpec{sk::Backend::Type::OpenMP});
            } catch (const bs_compat::Error&) {
                refused = 1;
            }
            map.setBackend(sk::BackendSpec{sk::Backend::Type::CUDA});
            for (int run = 0; run < bs_compat::kNumRuns; ++run) map(pv, sv, kv, rv, vv, tv, tyv, data);
            pv.flush();
            up_s = sv.uploads();
            down_p = pv.downloads();
        }
    } catch (const bs_compat::Error& e) {
        fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    FILE* o = fopen(argv[2], "w");
    if (!o) return 2;
    for (int i = 0; i < n; ++i)
        fprintf(o, "%.9g %.9g %.9g %.9g %.9g %.9g %.9g\n", p_map[i], p_thread[i], p_batch[i],
                i < 64 ? p_scalar[i] : 0.0f, i < 64 ? cndf[i] : 0.0f, p_pf[i], p_sk[i]);
    fprintf(o, "# uploads_sptprice %d downloads_prices %d cpu_backend_refused %d\n", up_s, down_p,
            refused);
    fclose(o);
    return 0;
}
