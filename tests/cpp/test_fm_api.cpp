// Drives the C++ drop-in (include/fm/sparse_matrix.h) the way the reference's
// own drivers do (matrix/test/test-2d_multiply.cpp:94-139,
// matrix/test/rand_mat_gen.cpp:178-199): build a 2D matrix from an edge
// list, multiply by a dense matrix, check the known answer A*[0..n-1].
// Prints "OK" on success; exit status 0/1.
#include <cstdio>
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <string>
#include <unistd.h>
#include <vector>

#include "fm/eigensolver.h"
#include "fm/sparse_matrix.h"

int main()
{
    if (!fm::init_flash_matrix(0)) return 1;
    const uint32_t n = 5000;
    std::vector<uint32_t> src, dst;
    for (uint32_t i = 0; i < n; i++)
        for (uint32_t j = 1; j <= 3; j++) {
            src.push_back(i);
            dst.push_back((i * 7 + j * 131) % n);
        }
    uint32_t *ds, *dd;
    hipMalloc(&ds, src.size() * 4);
    hipMalloc(&dd, dst.size() * 4);
    hipMemcpy(ds, src.data(), src.size() * 4, hipMemcpyHostToDevice);
    hipMemcpy(dd, dst.data(), dst.size() * 4, hipMemcpyHostToDevice);
    fm::sparse_matrix::ptr A = fm::sparse_matrix::create_2d_matrix(ds, dd, src.size(),
                                                                   fm::block_2d_size(256, 1024), false, n);
    if (!A || A->get_num_rows() != n || A->is_symmetric()) return 1;
    std::vector<long> x(n);
    for (uint32_t i = 0; i < n; i++) x[i] = i;
    fm::dense_matrix::ptr X = fm::dense_matrix::create_from_host(x.data(), n, 1, FMX_I64);
    fm::dense_matrix::ptr Y = A->multiply(X);
    fm::dense_matrix::ptr Yt = A->transpose()->multiply(X);
    if (!Y || !Yt) return 1;
    std::vector<long> y(n), yt(n), want(n, 0), want_t(n, 0);
    Y->copy_to_host(y.data());
    Yt->copy_to_host(yt.data());
    for (size_t e = 0; e < src.size(); e++) {
        if (src[e] == dst[e]) continue;
        bool dup = false;  // dedup: each (src, dst) once
        for (size_t f = 0; f < e; f++) dup |= (src[f] == src[e] && dst[f] == dst[e]);
        if (dup) continue;
        want[src[e]] += dst[e];
        want_t[dst[e]] += src[e];
    }
    for (uint32_t i = 0; i < n; i++)
        if (y[i] != want[i] || yt[i] != want_t[i]) {
            printf("mismatch at %u: %ld/%ld %ld/%ld\n", i, y[i], want[i], yt[i], want_t[i]);
            return 1;
        }
    // the .mat / .mat_idx pair: dump, then load the way the reference's
    // drivers do (matrix/test/test-2d_multiply.cpp:147-161)
    {
        char dir[] = "/tmp/fmx_cpp_XXXXXX";
        if (!mkdtemp(dir)) return 1;
        const std::string mat = std::string(dir) + "/a.mat", idx = std::string(dir) + "/a.mat_idx";
        if (!A->dump(mat, idx)) return 1;
        fm::SpM_2d_index::ptr index = fm::SpM_2d_index::load(idx);
        fm::sparse_matrix::ptr B = fm::sparse_matrix::create(index, fm::SpM_2d_storage::load(mat, index));
        if (!B || B->get_nnz() != A->get_nnz() || index->get_num_block_rows() != (n + 255) / 256) return 1;
        fm::dense_matrix::ptr Yb = B->multiply(X);
        std::vector<long> yb(n);
        Yb->copy_to_host(yb.data());
        if (yb != y) return 1;
        // out of core: the same file streamed through HBM in 4 KiB chunks
        // (sparse_matrix::create(index, safs::create_io_factory(...)),
        // matrix/test/test-2d_multiply.cpp:156-158)
        fm::sparse_matrix::ptr C =
            fm::sparse_matrix::create(index, safs::create_io_factory(mat, safs::REMOTE_ACCESS, 4096));
        if (!C || C->is_in_mem() || !B->is_in_mem() || C->get_nnz() != A->get_nnz()) return 1;
        std::vector<long> yc(n);
        C->multiply(X)->copy_to_host(yc.data());
        if (yc != y) return 1;
        bool io_threw = false;
        try {
            fm::SpM_2d_index::load(std::string(dir) + "/missing.mat_idx");
        } catch (const fm::io_exception &) {
            io_threw = true;
        }
        if (!io_threw) return 1;
        remove(mat.c_str());
        remove(idx.c_str());
        rmdir(dir);
    }
    // PageRank over A and its transpose (matrix/libmatrix-algs/PageRank.cpp),
    // against a host power iteration of the same recurrence
    {
        fm::dense_matrix::ptr pr = fm::PageRank(A, 100, 0.85f);
        if (!pr) return 1;
        std::vector<float> g(n);
        pr->copy_to_host(g.data());
        std::vector<double> outdeg(n, 0), p(n, 1.0 / n), in(n), out(n);
        std::vector<std::pair<uint32_t, uint32_t>> edges;
        for (size_t e = 0; e < src.size(); e++) {
            if (src[e] == dst[e]) continue;
            bool dup = false;
            for (size_t f = 0; f < e && !dup; f++) dup = (src[f] == src[e] && dst[f] == dst[e]);
            if (!dup) edges.push_back({src[e], dst[e]});
        }
        for (auto &e : edges) outdeg[e.first] += 1;
        for (int it = 0; it < 100; it++) {
            for (uint32_t i = 0; i < n; i++) in[i] = p[i] / std::max(1.0, outdeg[i]);
            std::fill(out.begin(), out.end(), 0.0);
            for (auto &e : edges) out[e.second] += in[e.first];
            for (uint32_t i = 0; i < n; i++) p[i] = 0.85 * out[i] + 0.15 / n;
        }
        for (uint32_t i = 0; i < n; i++)
            if (std::abs(g[i] - p[i]) > 1e-2 / n) {  // stopped once every step < 0.001/n
                printf("pagerank mismatch at %u: %g vs %g\n", i, g[i], p[i]);
                return 1;
            }
    }
    // NMF over A and its transpose (matrix/libmatrix-algs/NMF.cpp): positive
    // factors, and more iterations fit A no worse (Lee-Seung monotonicity),
    // measured here densely on the host
    {
        auto fit = [&](size_t iters) {
            auto f = fm::alg::NMF(A, 4, iters);
            if (!f.first || !f.second) return -1.0;
            std::vector<float> w(n * 4), h(n * 4);
            f.first->copy_to_host(w.data());
            f.second->copy_to_host(h.data());
            std::vector<double> a((size_t)n * n, 0.0);
            for (size_t e = 0; e < src.size(); e++)
                if (src[e] != dst[e]) a[(size_t)src[e] * n + dst[e]] = 1.0;
            double d = 0;
            for (uint32_t i = 0; i < n; i++)
                for (uint32_t j = 0; j < n; j++) {
                    double p = 0;
                    for (int c = 0; c < 4; c++) {
                        if (!(w[i * 4 + c] > 0) || !(h[j * 4 + c] > 0)) return -1.0;
                        p += (double)w[i * 4 + c] * h[j * 4 + c];
                    }
                    d += (a[(size_t)i * n + j] - p) * (a[(size_t)i * n + j] - p);
                }
            return d;
        };
        const double d1 = fit(1), d20 = fit(20);
        if (d1 < 0 || d20 < 0 || d20 > d1) {
            printf("NMF: %g after 1 iteration, %g after 20\n", d1, d20);
            return 1;
        }
    }
    // column-major input and output (L_COL; test-eigen.cpp:115-118), an input
    // in host memory streamed through HBM in column panels (the panel branch
    // of multiply(dense_matrix::ptr, mem_size), sparse_matrix.cpp:796-830),
    // and a creator with row scales (multiply(in, out, creator),
    // sparse_matrix.cpp:709-741): all equal the row-major product exactly on
    // integer-valued FP64 data
    {
        const size_t k = 3;
        std::vector<double> xr(n * k), xc(n * k);
        for (uint32_t i = 0; i < n; i++)
            for (size_t j = 0; j < k; j++) {
                const double v = (double)((i * 31 + j * 7) % 97);
                xr[i * k + j] = v;
                xc[j * n + i] = v;
            }
        fm::dense_matrix::ptr Xr = fm::dense_matrix::create_from_host(xr.data(), n, k, FMX_F64);
        fm::dense_matrix::ptr Xc = fm::dense_matrix::create(n, k, FMX_F64, fm::L_COL);
        fm::dense_matrix::ptr Yc = fm::dense_matrix::create(n, k, FMX_F64, fm::L_COL);
        if (!Xr || !Xc || !Yc || Xc->store_layout() != fm::L_COL) return 1;
        hipMemcpy(Xc->get_raw_arr(), xc.data(), n * k * 8, hipMemcpyHostToDevice);
        fm::dense_matrix::ptr Yr = A->multiply(Xr);
        if (!Yr || !A->multiply(*Xc, *Yc)) return 1;
        std::vector<double> yr(n * k), yc(n * k), yh(n * k), ys(n * k);
        Yr->copy_to_host(yr.data());
        Yc->copy_to_host(yc.data());
        if (yr != yc) return 1;
        fm::dense_matrix::ptr Xh = fm::dense_matrix::create_host(n, k, FMX_F64);
        if (!Xh || Xh->is_in_mem()) return 1;
        memcpy(Xh->get_raw_arr(), xr.data(), n * k * 8);
        fm::dense_matrix::ptr Yh = A->multiply(Xh, (size_t)2 * n * 8);  // one column per panel
        if (!Yh || Yh->is_in_mem()) return 1;
        Yh->copy_to_host(yh.data());
        if (yh != yr) return 1;
        if (A->multiply(Xh, 8)) return 1;  // the memory is too small for the SpMM
        std::vector<double> half(n, 0.5), two(n, 2.0);
        double *d_half, *d_two;
        hipMalloc(&d_half, n * 8);
        hipMalloc(&d_two, n * 8);
        hipMemcpy(d_half, half.data(), n * 8, hipMemcpyHostToDevice);
        hipMemcpy(d_two, two.data(), n * 8, hipMemcpyHostToDevice);
        fm::dense_matrix::ptr Ys = fm::dense_matrix::create(n, k, FMX_F64);
        if (!Ys || !A->multiply(*Xr, *Ys, fm::detail::spmm_creator::create(d_two, d_half, true))) return 1;
        Ys->copy_to_host(ys.data());
        if (ys != yr) return 1;  // 0.5 * A * (2 X) = A X exactly
        hipFree(d_half);
        hipFree(d_two);
    }
    // the eigensolver front end (matrix/eigensolver/eigensolver.h) with the
    // reference driver's operators (matrix/test/test-eigen.cpp): each pair
    // returned satisfies ||run(x) - theta x|| <= tol |theta| under the
    // operator's own FP64 run(), and the top eigenvalue of D^-1/2 S D^-1/2
    // is 1.  The graph: vertex i has 1 + i % 8 out-edges, to targets skewed
    // towards low ids, so that the top eigenvalues are well separated (the
    // graph above is nearly regular: its top eigenvalues of A^T A agree to
    // 1e-6)
    {
        std::vector<uint32_t> s2, d2;
        auto h = [](uint64_t x) { return (uint32_t)((x * 2654435761ull) & 0xFFFFFFFFull); };
        for (uint32_t i = 0; i < n; i++)
            for (uint32_t j = 0; j <= i % 8; j++) {
                const uint64_t x = (uint64_t)i * 16 + j;
                s2.push_back(i);
                d2.push_back(h(x) % (1 + h(x + 7) % n));
            }
        uint32_t *ds2, *dd2;
        if (hipMalloc(&ds2, s2.size() * 4) != hipSuccess || hipMalloc(&dd2, d2.size() * 4) != hipSuccess) return 1;
        if (hipMemcpy(ds2, s2.data(), s2.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(dd2, d2.data(), d2.size() * 4, hipMemcpyHostToDevice) != hipSuccess)
            return 1;
        fm::sparse_matrix::ptr A = fm::sparse_matrix::create_2d_matrix(ds2, dd2, s2.size(),
                                                                       fm::block_2d_size(256, 1024), false, n);
        fm::sparse_matrix::ptr S = fm::sparse_matrix::create_2d_matrix(
            ds2, dd2, s2.size(), fm::block_2d_size(256, 1024), true, n, FMX_BUILD_DEDUP | FMX_BUILD_SYMMETRIZE);
        (void)hipFree(ds2);
        (void)hipFree(dd2);
        if (!A || !A->transpose() || !S || !S->is_symmetric()) return 1;
        const int nev = 4;
        const double tol = 1e-8;
        auto check = [&](fm::eigen::spm_function *op, const fm::eigen::spm_function &again, const char *name) {
            fm::eigen::eigen_options opts;
            if (!opts.init(nev)) return false;
            opts.tol = tol;
            opts.precision = FMX_EIG_FP64;
            fm::eigen::eigen_res res = fm::eigen::compute_eigen(op, true, opts);
            if (!res.converged || (int)res.vals.size() != nev || !res.vecs) {
                printf("%s: no convergence\n", name);
                return false;
            }
            const size_t rows = res.vecs->get_num_rows();
            std::vector<double> v(rows * nev), y(rows * nev);
            res.vecs->copy_to_host(v.data());
            fm::dense_matrix::ptr x = fm::dense_matrix::create_from_host(v.data(), rows, nev, FMX_F64);
            fm::dense_matrix::ptr ny = again.run(x);
            if (!ny || x) return false;  // run() releases its input, as the reference's does
            ny->copy_to_host(y.data());
            for (int i = 0; i < nev; i++) {
                double r = 0.0;
                for (size_t j = 0; j < rows; j++) {
                    const double d = y[j * nev + i] - res.vals[i] * v[j * nev + i];
                    r += d * d;
                }
                if (std::sqrt(r) > 1.01 * tol * std::fabs(res.vals[i]) || (i && res.vals[i] > res.vals[i - 1])) {
                    printf("%s: pair %d residual %.3e, value %.12f\n", name, i, std::sqrt(r) / std::fabs(res.vals[i]),
                           res.vals[i]);
                    return false;
                }
            }
            if (std::string(name) == "NA_eigen" && std::fabs(res.vals[0] - 1.0) > 1e-8) return false;
            return true;
        };
        if (!check(new fm::eigen::NA_eigen_Operator(S), fm::eigen::NA_eigen_Operator(S), "NA_eigen")) return 1;
        if (!check(new fm::eigen::eigen_Operator(S), fm::eigen::eigen_Operator(S), "eigen")) return 1;
        if (!check(new fm::eigen::SVD_Operator(A), fm::eigen::SVD_Operator(A), "SVD")) return 1;
        if (!check(new fm::eigen::SVD_Operator(S), fm::eigen::SVD_Operator(S), "SVD (symmetric)")) return 1;
        fm::eigen::eigen_options lm;
        lm.init(nev);
        lm.which = "LM";
        if (!fm::eigen::compute_eigen(new fm::eigen::eigen_Operator(S), true, lm).vals.empty()) return 1;
    }
    // error convention: a dimension mismatch is false + log, not an exception
    fm::dense_matrix::ptr bad = fm::dense_matrix::create(n + 1, 1, FMX_I64);
    if (A->multiply(*bad, *Y)) return 1;
    bool threw = false;
    try {
        std::vector<char> junk(8192, 0);
        std::vector<int64_t> offs = {4096, 8192};
        fm::sparse_matrix::create(junk, offs);
    } catch (const fm::wrong_format &) {
        threw = true;
    }
    if (!threw) return 1;
    hipFree(ds);
    hipFree(dd);
    fm::destroy_flash_matrix();
    printf("OK\n");
    return 0;
}
