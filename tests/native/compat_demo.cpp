// Drives include/bs_compat.hpp the way the reference's drivers use their
// entry points: the FastFlow map (run/wait, 100 passes), bs_thread over bound
// globals, the SSE pointer-batch form and the per-option forms.  Writes one
// "%.9g" line per price of each form, for tests/test_gpu_tools.py.
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
    } catch (const bs_compat::Error& e) {
        fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    FILE* o = fopen(argv[2], "w");
    if (!o) return 2;
    for (int i = 0; i < n; ++i)
        fprintf(o, "%.9g %.9g %.9g %.9g %.9g\n", p_map[i], p_thread[i], p_batch[i],
                i < 64 ? p_scalar[i] : 0.0f, i < 64 ? cndf[i] : 0.0f);
    fclose(o);
    return 0;
}
