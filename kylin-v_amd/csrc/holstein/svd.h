// Bond SVD with truncation, FP64 on the device.
// Follows svd_impl<R1,R2> (holstein/svd.h:296-440); the sector structure and the
// truncation rule are in svd_rule.h.
//
// Per charge sector three FP64 routes (sectors whose smaller side is below iso_min, and
// zero-keeping truncations that cannot take a Gram route, use the first):
//  * exact: rocSOLVER one-sided Jacobi (zgesvdj), or zgesvd when cutoff < 0 keeps zero
//    singular values (their vectors must still be an orthonormal completion);
//  * smaller-side Gram (when the isometry is the larger side): G on the smaller side by
//    the MFMA GEMM, Hermitian eigen-decomposition (batched zhetrd, dstedc per sector,
//    eigenvectors back-transformed only for the kept values), the isometry as M V / s by
//    one more GEMM.  The eigenproblem of the two 2048-side Gram matrices of a 32-site,
//    D=256 bond takes ~0.07 s on MI355X against ~2 s per sector for the Jacobi SVD.
//    M V / s loses orthonormality like eps (s_max / s)^2, so a sector that keeps values
//    below guard * s_max is redone on the isometric side;
//  * isometric-side Gram (when the isometry is the smaller or equal side):
//    G = M M^H when the left factor is the isometry (S goes right), M^H M when the right
//    one is; the eigenvectors of G are the isometry itself, orthonormal to machine
//    precision, and the other factor is one GEMM (U^H M or M V) that already carries S,
//    so no division by small singular values.  Eigenvalues of G are only accurate to
//    ~eps lambda_max, so when the truncation reaches below tail * s_max the tail
//    eigenvectors (an accurate subspace as a whole) are refined by a second Gram pass on
//    M restricted to that subspace, whose norm is below tail * s_max: the tail's singular
//    values then carry absolute error ~eps (tail s_max)^2 / s.
// Layout of the factors: U_M (rows x k) and V_M^H (k x cols) row-major.  rocSOLVER is
// column-major, so the exact route hands it M^T (cols x rows): with M^T = U' S V'^H,
// M = conj(V') S U'^T, so rocSOLVER's V^H output read row-major (ld k) is U_M and its U
// output read row-major (ld cols) is V_M^H, with no extra transposes.
#pragma once

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "svd_rule.h"
#include "tensor.h"

namespace kv {

// rocBLAS initialises its Tensile host state lazily on the first handle; concurrent first
// calls from several threads fail ("Could not initialize Tensile host"), so handle
// creation is serialised process-wide.
inline std::mutex& solver_init_mutex() {
    static std::mutex m;
    return m;
}

struct Solver {
    rocblas_handle h = nullptr;
    // A second handle on a stream of its own (default priority), made on first use: the
    // tridiagonal eigensolves of a block-route bond's two sectors run side by side.
    rocblas_handle h2 = nullptr;
    hipStream_t s2 = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    explicit Solver(Device& dev) {
        std::lock_guard<std::mutex> g(solver_init_mutex());
        if (rocblas_create_handle(&h) != rocblas_status_success)
            throw Error(ERR_SOLVER, "rocblas_create_handle failed");
        if (rocblas_set_stream(h, dev.stream) != rocblas_status_success)
            throw Error(ERR_SOLVER, "rocblas_set_stream failed");
    }
    rocblas_handle second() {
        if (h2) return h2;
        std::lock_guard<std::mutex> g(solver_init_mutex());
        KV_HIP(hipStreamCreateWithFlags(&s2, hipStreamNonBlocking));
        KV_HIP(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
        KV_HIP(hipEventCreateWithFlags(&join, hipEventDisableTiming));
        if (rocblas_create_handle(&h2) != rocblas_status_success)
            throw Error(ERR_SOLVER, "rocblas_create_handle failed");
        if (rocblas_set_stream(h2, s2) != rocblas_status_success)
            throw Error(ERR_SOLVER, "rocblas_set_stream failed");
        return h2;
    }
    ~Solver() {
        if (h) rocblas_destroy_handle(h);
        if (h2) rocblas_destroy_handle(h2);
        if (fork) (void)hipEventDestroy(fork);
        if (join) (void)hipEventDestroy(join);
        if (s2) (void)hipStreamDestroy(s2);
    }
    Solver(const Solver&) = delete;
    Solver& operator=(const Solver&) = delete;
};

struct SvdResult {
    std::unique_ptr<Tensor> left, right;
    int kept = 0;
    int total = 0;                           // singular values before truncation
    std::vector<std::vector<double>> svals;  // per sector pair, kept values
    int gram_sectors = 0, exact_sectors = 0, refined_sectors = 0;
    int block_sectors = 0, block_fallbacks = 0;  // block subspace route: kept, redone in full
    std::map<int, int> kept_by_q;                 // kept dimension per row-sector charge
    std::string detail;  // per sector: rows x cols, route, s_max, last kept value
};

// Side from which the split eigensolver replaces zheevd by default.  Below it the
// back-transformation of all vectors inside zheevd is cheap, and the split route's extra
// launches and host synchronisations cost more than they save: with the split route at
// every side, 8 concurrent chains of the batching workload ran 7% slower.  From 512 up the
// split route wins: rocSOLVER's column-by-column reduction dominates zheevd there (the
// 32-site chain's edge bonds, 512-wide Gram matrices, and the batching chains' largest
// sectors).
constexpr int kSplitEigMin = 512;
// The side from which the split route is taken (KV_SPLIT_MIN overrides kSplitEigMin).
inline int split_eig_min() {
    static const int v = [] {
        const char* e = std::getenv("KV_SPLIT_MIN");
        const int x = e ? std::atoi(e) : 0;
        return x >= 2 ? x : kSplitEigMin;
    }();
    return v;
}
enum EigMode { EIG_AUTO = 0, EIG_SPLIT = 1, EIG_HEEVD = 2 };

struct SvdOptions {
    bool block = true;      // block subspace route for sectors far larger than the kept dimension
    double guard = 1e-4;    // smaller-side Gram vectors only if s_kept_min >= guard * s_max
    int iso_min = 64;       // smaller side from which the Gram routes are used
    double block_min = 3.0; // block route for sectors whose smaller side is >= block_min x the block
                            // (at 2x, 512 x 2048 sectors keeping half their values exhausted
                            // the power steps and fell back after the work was spent)
    double tail = 1e-5;     // isometric route: values below tail * s_max get a second pass
    bool force_exact = false;
    int eig_mode = EIG_AUTO;  // Gram eigensolver: split from kSplitEigMin, or forced
    bool roc_trd = false;     // split route: rocSOLVER zhetrd instead of hetrd_lower
    bool qr_reduce = true;    // Householder-reduce long isometric sides before their Gram route
    static SvdOptions from_env() {
        SvdOptions o;
        if (const char* e = std::getenv("KV_QR_REDUCE")) o.qr_reduce = std::string(e) != "0";
        if (const char* e = std::getenv("KV_TRD")) o.roc_trd = std::string(e) == "rocsolver";
        if (const char* e = std::getenv("KV_SVD")) {
            o.force_exact = std::string(e) == "exact";
            o.block = std::string(e) != "full" && !o.force_exact;
        }
        if (const char* e = std::getenv("KV_EIG"))
            o.eig_mode = std::string(e) == "heevd" ? EIG_HEEVD : std::string(e) == "split" ? EIG_SPLIT : EIG_AUTO;
        if (const char* g = std::getenv("KV_GRAM_GUARD")) o.guard = std::atof(g);
        if (const char* m = std::getenv("KV_ISO_MIN")) o.iso_min = std::atoi(m);
        if (const char* m = std::getenv("KV_BLOCK_MIN")) o.block_min = std::atof(m);
        if (const char* t = std::getenv("KV_ISO_TAIL")) o.tail = std::atof(t);
        return o;
    }
};

// C (m x n, ld ldc) = op(A) op(B) for one dense complex problem on the grouped GEMM.
// upper_only: only the 64 x 64 tiles on or above the diagonal (Hermitian products whose
// consumer reads one triangle); the rest of C is left as it was.
inline void gemm_dense(Device& D, char opA, char opB, const double2* A, int lda, const double2* B,
                       int ldb, double2* C, int ldc, int m, int n, int k, bool upper_only = false) {
    GemmGroup g{};
    g.m = m;
    g.n = n;
    g.ldc = ldc;
    g.seg_begin = 0;
    g.seg_count = 1;
    GemmSeg sg{};
    sg.k = k;
    sg.lda = lda;
    sg.ldb = ldb;
    std::vector<GemmGroup> gs{g};
    std::vector<GemmSeg> ss{sg};
    std::vector<GemmTile> ts;
    for (int a = 0; a < (m + GEMM_BM - 1) / GEMM_BM; ++a)
        for (int b = 0; b < (n + GEMM_BN - 1) / GEMM_BN; ++b)
            if (!upper_only || a * GEMM_BM < (b + 1) * GEMM_BN) ts.push_back(GemmTile{0, a, b, 0});
    PackedUpload pk = upload_packed(D, gs, ss, ts);
    launch_gemm<double2>(D.stream, op_code(opA), op_code(opB), A, B, C, pk.at<GemmTile>(2),
                         (int)ts.size(), pk.at<GemmGroup>(0), pk.at<GemmSeg>(1), 1.0, 0.0);
}

struct SectorSvd {
    int64_t nr = 0, nc = 0;
    int k = 0;            // number of singular values available (Gram: side of G)
    bool gram = false;    // values from the Gram eigenproblem, vectors not yet formed
    bool iso = false;     // isometric-side Gram: the eigenvectors are the isometry
    bool refined = false; // isometric route: tail already refined
    bool colside = true;  // Gram of M^H M (else M M^H)
    bool s_absorbed = false;  // the non-isometric factor already carries S
    double gscale = 1.0;      // G was divided by this (|M|_F^2) before the eigensolver
    DevBuf M, G, S, U, Vh;  // U: nr x ldU, Vh: k x nc
    int ldU = 0;
    std::vector<double> s;  // descending
    // Gram eigenvectors: rows [ev_lo, k) of the ascending eigenvector matrix, row j - ev_lo of
    // Ev (ld k) read row-major is v_j^H.  The split solver leaves G holding its Householder
    // reflectors, tau their scalars and Zt the tridiagonal eigenvectors (real, column-major),
    // from which gram_eigvecs forms the rows on demand.
    DevBuf Ev, tau, Zt;
    int ev_lo = 0;
    // block subspace route: Z (isometric side x kc) the leading Ritz vectors, W (other side
    // x kc) the sector matrix applied to them (it carries S)
    bool block = false;
    DevBuf Z, W;
    int kc = 0;
    double ceil = 0.0;  // every singular value beyond the certified ones is at most this
    // Householder reduction (qr_reduce): M holds the square triangular factor and Qf the
    // orthonormal one, row-major; the isometric side was `full` long
    bool reduced = false;
    int64_t full = 0;
    DevBuf Qf;
};

inline void exact_sector(Device& D, Solver& sol, SectorSvd& X, bool keep_zero) {
    const int k = (int)std::min(X.nr, X.nc);
    X.k = k;
    X.gram = false;
    X.ldU = k;
    // rocSOLVER overwrites its input: work on a copy so the Gram route can fall back
    DevBuf A(D.pool.get(), (size_t)(X.nr * X.nc) * sizeof(double2));
    KV_HIP(hipMemcpyAsync(A.get(), X.M.get(), (size_t)(X.nr * X.nc) * sizeof(double2),
                          hipMemcpyDeviceToDevice, D.stream));
    X.S = DevBuf(D.pool.get(), (size_t)k * sizeof(double));
    X.Vh = DevBuf(D.pool.get(), (size_t)(X.nc * k) * sizeof(double2));  // rocSOLVER U
    X.U = DevBuf(D.pool.get(), (size_t)(k * X.nr) * sizeof(double2));   // rocSOLVER V^H
    DevBuf E(D.pool.get(), (size_t)std::max(k, 1) * sizeof(double));
    DevBuf info(D.pool.get(), sizeof(rocblas_int));
    auto* a = reinterpret_cast<rocblas_double_complex*>(A.get());
    auto* u = reinterpret_cast<rocblas_double_complex*>(X.Vh.get());
    auto* v = reinterpret_cast<rocblas_double_complex*>(X.U.get());
    rocblas_status st;
    if (!keep_zero) {
        DevBuf res(D.pool.get(), sizeof(double)), nsw(D.pool.get(), sizeof(rocblas_int));
        st = rocsolver_zgesvdj(sol.h, rocblas_svect_singular, rocblas_svect_singular,
                               (rocblas_int)X.nc, (rocblas_int)X.nr, a, (rocblas_int)X.nc, 0.0,
                               res.as<double>(), 100, nsw.as<rocblas_int>(), X.S.as<double>(), u,
                               (rocblas_int)X.nc, v, (rocblas_int)k, info.as<rocblas_int>());
    } else {
        st = rocsolver_zgesvd(sol.h, rocblas_svect_singular, rocblas_svect_singular,
                              (rocblas_int)X.nc, (rocblas_int)X.nr, a, (rocblas_int)X.nc,
                              X.S.as<double>(), u, (rocblas_int)X.nc, v, (rocblas_int)k,
                              E.as<double>(), rocblas_outofplace, info.as<rocblas_int>());
    }
    KV_REQUIRE(st == rocblas_status_success, ERR_SOLVER, "rocSOLVER SVD failed");
    rocblas_int hinfo = 0;
    X.s.resize(k);
    KV_HIP(hipMemcpyAsync(&hinfo, info.get(), sizeof(rocblas_int), hipMemcpyDeviceToHost, D.stream));
    KV_HIP(hipMemcpyAsync(X.s.data(), X.S.get(), k * sizeof(double), hipMemcpyDeviceToHost, D.stream));
    D.sync();
    KV_REQUIRE(hinfo == 0, ERR_SOLVER, "Problems in SVD (did not converge)");
}

// Gram matrix of a sector (G = M^H M when cols <= rows, else M M^H; the isometric route
// takes the side given by colside).
inline void gram_prepare(Device& D, SectorSvd& X, bool iso = false, bool colside = true) {
    X.iso = iso;
    X.colside = iso ? colside : X.nc <= X.nr;
    const int n = (int)(X.colside ? X.nc : X.nr);
    X.k = n;
    X.gram = true;
    X.G = DevBuf(D.pool.get(), (size_t)n * n * sizeof(double2));
    const double2* m = X.M.as<double2>();
    // the eigensolvers read the lower triangle of the column-major view, i.e. the upper
    // triangle of this row-major buffer: only those tiles are computed, the rest is zero
    launch_zero(D.stream, X.G.get(), X.G.bytes());
    if (X.colside)  // G = M^H M
        gemm_dense(D, 'C', 'N', m, (int)X.nc, m, (int)X.nc, X.G.as<double2>(), n, n, n, (int)X.nr, true);
    else  // G = M M^H
        gemm_dense(D, 'N', 'C', m, (int)X.nc, m, (int)X.nc, X.G.as<double2>(), n, n, n, (int)X.nc, true);
    // G is handed to the eigensolver at unit trace (a sector of a normalised state can
    // have any small norm); the eigenvalues are scaled back after the solve
    DevBuf sc(D.pool.get(), 1024 * sizeof(double)), nr2(D.pool.get(), sizeof(double));
    launch_nrm2sq(D.stream, m, X.nr * X.nc, nr2.as<double>(), sc.as<double>());
    double f2 = 0;
    KV_HIP(hipMemcpyAsync(&f2, nr2.get(), sizeof(double), hipMemcpyDeviceToHost, D.stream));
    D.sync();
    X.gscale = f2 > 0 ? f2 : 1.0;
    launch_zaxpby(D.stream, X.G.as<double2>(), X.G.as<double2>(), nullptr,
                  make_double2(1.0 / X.gscale, 0.0), make_double2(0.0, 0.0), (int64_t)n * n);
}

// Householder reduction of a sector whose isometric side is the longer one.  The row-major
// buffer of M (nr x nc) is the column-major nc x nr matrix M^T.  Column-side isometry
// (nc > nr): M^T = Q R (zgeqrf), so M = R^T Q^T and the sector continues as the nr x nr
// lower-triangular R^T, with V^H = V'^H Q^T at the end.  Row-side isometry (nr > nc):
// M^T = L Q (zgelqf), M = Q^T L^T, the sector continues as the nc x nc upper-triangular L^T,
// with U = Q^T U'.  Both triangles sit where the row-major M had them (element (i, j) at
// i * nc + j), and the explicit Q (zungqr / zunglq) read row-major is Q^T.  The isometric
// Gram matrix is then square in the shorter side instead of the longer one.
inline void qr_reduce(Device& D, Solver& sol, SectorSvd& X, bool colside) {
    const int r = (int)X.nr, c = (int)X.nc;
    const int m = colside ? r : c;
    DevBuf A(D.pool.get(), X.M.bytes()), tau(D.pool.get(), (size_t)m * sizeof(double2));
    KV_HIP(hipMemcpyAsync(A.get(), X.M.get(), X.M.bytes(), hipMemcpyDeviceToDevice, D.stream));
    auto* a = reinterpret_cast<rocblas_double_complex*>(A.get());
    auto* t = reinterpret_cast<rocblas_double_complex*>(tau.get());
    auto ok = [](rocblas_status st, const char* what) {
        KV_REQUIRE(st == rocblas_status_success, ERR_SOLVER, std::string(what) + " failed");
    };
    if (colside)
        ok(rocsolver_zgeqrf(sol.h, c, r, a, c, t), "rocsolver_zgeqrf");
    else
        ok(rocsolver_zgelqf(sol.h, c, r, a, c, t), "rocsolver_zgelqf");
    DevBuf T(D.pool.get(), (size_t)m * m * sizeof(double2));
    launch_tri_copy(D.stream, A.as<double2>(), c, T.as<double2>(), m, colside);
    if (colside)
        ok(rocsolver_zungqr(sol.h, c, r, r, a, c, t), "rocsolver_zungqr");
    else
        ok(rocsolver_zunglq(sol.h, c, r, c, a, c, t), "rocsolver_zunglq");
    X.reduced = true;
    X.full = colside ? c : r;
    X.Qf = std::move(A);
    X.M = std::move(T);
    X.nr = X.nc = m;
}

// Undo qr_reduce on the formed factors: V^H (k x m) <- V^H Q^T (k x nc), or U (m x ldU)
// <- Q^T U (nr x ldU).
inline void qr_expand(Device& D, SectorSvd& X, bool colside, int nk) {
    const int m = (int)X.nr;
    const int n = (int)X.full;
    if (colside) {
        DevBuf Vh(D.pool.get(), (size_t)nk * n * sizeof(double2));
        gemm_dense(D, 'N', 'N', X.Vh.as<double2>(), m, X.Qf.as<double2>(), n, Vh.as<double2>(), n, nk, n, m);
        X.Vh = std::move(Vh);
        X.nc = n;
    } else {
        DevBuf U(D.pool.get(), (size_t)n * X.ldU * sizeof(double2));
        gemm_dense(D, 'N', 'N', X.Qf.as<double2>(), m, X.U.as<double2>(), X.ldU, U.as<double2>(), X.ldU, n,
                   X.ldU, m);
        X.U = std::move(U);
        X.nr = n;
    }
    X.Qf = DevBuf();
    X.reduced = false;
}

// Reflectors per block of the back-transformation.
constexpr int kReflectorBlock = 128;

// Sum of squares of `count` doubles on the device (NaN or Inf if any is).
inline double dev_sumsq(Device& D, const double* p, int64_t count) {
    DevBuf sc(D.pool.get(), 2048 * sizeof(double)), out(D.pool.get(), sizeof(double));
    double tail = 0.0, v = 0.0;
    if (count >= 2) launch_nrm2sq(D.stream, reinterpret_cast<const double2*>(p), count / 2, out.as<double>(),
                                  sc.as<double>());
    if (count % 2)
        KV_HIP(hipMemcpyAsync(&tail, p + count - 1, sizeof(double), hipMemcpyDeviceToHost, D.stream));
    if (count >= 2) KV_HIP(hipMemcpyAsync(&v, out.get(), sizeof(double), hipMemcpyDeviceToHost, D.stream));
    D.sync();
    return v + tail * tail;
}

// Batched Householder tridiagonalisation of `batch` column-major N x N Hermitian matrices
// at stride N * N, of which the lower triangle is read.  On exit the storage is LAPACK
// zhetrd('L')'s: d = diag(T), e = subdiag(T), tau and the reflector tails below the
// subdiagonal (the upper triangle is overwritten).  Per column two launches (trd_column:
// the column brought up to date by the panel and reflected; trd_matvec: the trailing
// matrix times the reflector, as column dot products over the fully stored matrix), and per
// panel of kTrdPanel columns the rank-2 update of the trailing matrix, both triangles, as two
// rocBLAS GEMMs.  rocSOLVER's one-stage zhetrd spends five launches per column.
inline void hetrd_lower(Device& D, Solver& sol, double2* A, int N, int batch, double* d, double* e,
                        double2* tau) {
    const int nb = kTrdPanel;
    DevBuf VW(D.pool.get(), (size_t)N * 2 * nb * batch * sizeof(double2));
    DevBuf y(D.pool.get(), (size_t)N * batch * sizeof(double2));
    DevBuf q(D.pool.get(), (size_t)2 * nb * batch * sizeof(double2));
    DevBuf part(D.pool.get(), (size_t)trd_column_blocks(N, 0) * batch * sizeof(double2));
    DevBuf cnt(D.pool.get(), (size_t)batch * sizeof(unsigned));
    launch_zero(D.stream, cnt.get(), cnt.bytes());
    launch_herm_fill(D.stream, A, N, batch);
    double2* vw = VW.as<double2>();
    auto* ac = reinterpret_cast<rocblas_double_complex*>(A);
    auto* vc = reinterpret_cast<rocblas_double_complex*>(vw);
    const rocblas_double_complex one(1.0, 0.0), mone(-1.0, 0.0);
    const rocblas_stride sp = (rocblas_stride)N * 2 * nb, sa = (rocblas_stride)N * N;
    for (int p = 0; p < N; p += nb) {
        const int pw = std::min(nb, N - p);
        for (int i = 0; i < pw; ++i) {
            const int j = p + i;
            launch_trd_column(D.stream, A, vw, y.as<double2>(), q.as<double2>(), d, e, tau, part.as<double2>(),
                              cnt.as<unsigned>(), N, j, i, nb, i > 0, true, batch);
            if (j <= N - 2)
                launch_trd_matvec(D.stream, A, vw, y.as<double2>(), q.as<double2>(), N, j + 1, i, nb, batch);
        }
        const int t = p + pw, m = N - t;
        if (m <= 0) break;  // the panel ended on the last column
        // the last column's companion, then A(t:, t:) -= V W^H + W V^H
        launch_trd_column(D.stream, A, vw, y.as<double2>(), q.as<double2>(), d, e, tau, part.as<double2>(),
                          cnt.as<unsigned>(), N, t, pw, nb, true, false, batch);
        for (int h = 0; h < 2; ++h) {
            const rocblas_double_complex* a = vc + (h ? (size_t)nb * N : 0) + t;
            const rocblas_double_complex* b = vc + (h ? 0 : (size_t)nb * N) + t;
            const rocblas_status st = rocblas_zgemm_strided_batched(
                sol.h, rocblas_operation_none, rocblas_operation_conjugate_transpose, m, m, pw, &mone, a, N, sp, b,
                N, sp, &one, ac + (size_t)t * N + t, N, sa, batch);
            KV_REQUIRE(st == rocblas_status_success, ERR_SOLVER, "rocBLAS zgemm failed");
        }
    }
}

// Eigen-decomposition of the Gram matrices of one or more sectors.  The row-major
// Hermitian buffer read column-major is conj(G): eigenvalues are equal and the
// eigenvectors are conj(v_j), so row j - ev_lo of Ev read row-major is v_j^H (ascending).
// Tridiagonal route (sides >= kSplitEigMin by default, or sectors of different sides): one
// batched zhetrd for all sectors (its cost is per column, not per matrix, so the sectors of
// a bond cost about as much as the largest), every sector padded to the largest side N as
// diag(G, -I) (G is positive semidefinite, so the padding stays exactly decoupled and each
// sector's reflectors and tridiagonal are those of G); then dstedc per sector at its own
// side, and the vectors formed later, only for the rows the truncation needs
// (gram_eigvecs).  zheevd (all vectors at once, batched) for equal sides below
// kSplitEigMin.  (zheevd cannot take the padded matrices: on the split tridiagonal its
// divide and conquer returns NaN eigenvectors.  dstedc of a pair on two streams measured
// slower in the sweep than one after the other.)
inline void gram_solve(Device& D, Solver& sol, std::vector<SectorSvd*> xs, int mode = EIG_AUTO,
                       bool roc_trd = false, bool overlap = false) {
    const int nb = (int)xs.size();
    int N = 0;
    bool same = true;
    for (auto* x : xs) {
        same = same && (N == 0 || x->k == N);
        N = std::max(N, x->k);
    }
    const size_t NN = (size_t)N * N;
    const bool split = mode == EIG_SPLIT ? N >= 2
                       : mode == EIG_AUTO ? (N >= split_eig_min() || (!same && N >= 2))
                                          : false;
    KV_REQUIRE(split || same, ERR_INTERNAL, "zheevd needs sectors of one side");
    DevBuf W(D.pool.get(), (size_t)N * nb * sizeof(double)), E(D.pool.get(), (size_t)N * nb * sizeof(double));
    DevBuf info(D.pool.get(), sizeof(rocblas_int) * nb);
    // the solvers overwrite their input: they work on copies, so that a sector the divide
    // and conquer fails on can be redone from its Gram matrix (or tridiagonal form)
    DevBuf Gb(D.pool.get(), NN * nb * sizeof(double2)), taub;
    for (int b = 0; b < nb; ++b) {
        double2* slot = Gb.as<double2>() + b * NN;
        const int n = xs[b]->k;
        if (n == N) {
            KV_HIP(hipMemcpyAsync(slot, xs[b]->G.get(), NN * sizeof(double2), hipMemcpyDeviceToDevice,
                                  D.stream));
        } else {
            launch_zero(D.stream, slot, NN * sizeof(double2));
            KV_HIP(hipMemcpy2DAsync(slot, (size_t)N * sizeof(double2), xs[b]->G.get(), (size_t)n * sizeof(double2),
                                    (size_t)n * sizeof(double2), n, hipMemcpyDeviceToDevice, D.stream));
            launch_set_diag(D.stream, slot, N, n, N, -1.0);
        }
    }
    auto* gp = reinterpret_cast<rocblas_double_complex*>(Gb.get());
    std::vector<rocblas_int> hinfo(nb);
    auto read_info = [&] {
        KV_HIP(hipMemcpyAsync(hinfo.data(), info.get(), nb * sizeof(rocblas_int), hipMemcpyDeviceToHost,
                              D.stream));
        D.sync();
    };
    // rocSOLVER's divide and conquer (inside zheevd, and dstedc) can fail on a tridiagonal
    // that splits into exactly decoupled blocks (rank-deficient Gram matrices, e.g. of a
    // product state): it reports non-convergence, or returns NaN eigenvectors with info 0.
    // Such sectors are redone by implicit QR iteration (zheev / dsteqr).
    if (!split) {
        const int n = N;
        const rocblas_status st = rocsolver_zheevd_strided_batched(
            sol.h, rocblas_evect_original, rocblas_fill_lower, n, gp, n, (rocblas_stride)NN, W.as<double>(), n,
            E.as<double>(), n, info.as<rocblas_int>(), nb);
        KV_REQUIRE(st == rocblas_status_success, ERR_SOLVER, "rocsolver_zheevd failed");
        read_info();
        for (int b = 0; b < nb; ++b) {
            double2* slot = Gb.as<double2>() + b * NN;
            const bool ok = hinfo[b] == 0 && std::isfinite(dev_sumsq(D, reinterpret_cast<const double*>(slot),
                                                                      2 * (int64_t)NN));
            if (ok) continue;
            KV_HIP(hipMemcpyAsync(slot, xs[b]->G.get(), NN * sizeof(double2), hipMemcpyDeviceToDevice, D.stream));
            const rocblas_status s2 = rocsolver_zheev(sol.h, rocblas_evect_original, rocblas_fill_lower, n,
                                                      reinterpret_cast<rocblas_double_complex*>(slot), n,
                                                      W.as<double>() + (size_t)b * n, E.as<double>() + (size_t)b * n,
                                                      info.as<rocblas_int>() + b);
            KV_REQUIRE(s2 == rocblas_status_success, ERR_SOLVER, "rocsolver_zheev failed");
            read_info();
            KV_REQUIRE(hinfo[b] == 0 && std::isfinite(dev_sumsq(D, reinterpret_cast<const double*>(slot),
                                                                 2 * (int64_t)NN)),
                       ERR_SOLVER, "Hermitian eigensolver did not converge");
        }
        for (int b = 0; b < nb; ++b)
            KV_HIP(hipMemcpyAsync(xs[b]->G.get(), Gb.as<double2>() + b * NN, NN * sizeof(double2),
                                  hipMemcpyDeviceToDevice, D.stream));
    } else {
        taub = DevBuf(D.pool.get(), (size_t)N * nb * sizeof(double2));
        hipEvent_t trd_t0 = nullptr;
        if (roc_trd) {
            auto* tp = reinterpret_cast<rocblas_double_complex*>(taub.get());
            const rocblas_status st = rocsolver_zhetrd_strided_batched(
                sol.h, rocblas_fill_lower, N, gp, N, (rocblas_stride)NN, W.as<double>(), N,
                E.as<double>(), N, tp, N, nb);
            KV_REQUIRE(st == rocblas_status_success, ERR_SOLVER, "rocsolver_zhetrd failed");
        } else {
            trd_t0 = D.spans.mark(D.stream);
            if (N <= kTrdSmallMax) {  // one workgroup per matrix, no per-column launches
                launch_herm_fill(D.stream, Gb.as<double2>(), N, nb);
                launch_trd_small(D.stream, Gb.as<double2>(), N, W.as<double>(), E.as<double>(), taub.as<double2>(),
                                 nb);
            } else {
                hetrd_lower(D, sol, Gb.as<double2>(), N, nb, W.as<double>(), E.as<double>(), taub.as<double2>());
            }
            D.spans.close(trd_t0, D.stream, &D.trd.ms);
        }
        // the tridiagonal forms, kept for a redo (dstedc overwrites both arrays)
        DevBuf W0(D.pool.get(), W.bytes()), E0(D.pool.get(), E.bytes());
        KV_HIP(hipMemcpyAsync(W0.get(), W.get(), W.bytes(), hipMemcpyDeviceToDevice, D.stream));
        KV_HIP(hipMemcpyAsync(E0.get(), E.get(), E.bytes(), hipMemcpyDeviceToDevice, D.stream));
        // with `overlap`, a pair of sectors is solved on two streams (the second handle's
        // stream forks after everything queued so far and the context stream joins it)
        const bool two = overlap && nb == 2;
        if (two) {
            sol.second();
            KV_HIP(hipEventRecord(sol.fork, D.stream));
            KV_HIP(hipStreamWaitEvent(sol.s2, sol.fork, 0));
        }
        for (int b = 0; b < nb; ++b) xs[b]->Zt = DevBuf(D.pool.get(), (size_t)xs[b]->k * xs[b]->k * sizeof(double));
        for (int b = 0; b < nb; ++b) {
            const int n = xs[b]->k;
            const rocblas_status s2 = rocsolver_dstedc(
                two && b == 1 ? sol.h2 : sol.h, rocblas_evect_tridiagonal, n, W.as<double>() + (size_t)b * N,
                E.as<double>() + (size_t)b * N, xs[b]->Zt.as<double>(), n, info.as<rocblas_int>() + b);
            KV_REQUIRE(s2 == rocblas_status_success, ERR_SOLVER, "rocsolver_dstedc failed");
        }
        if (two) {
            KV_HIP(hipEventRecord(sol.join, sol.s2));
            KV_HIP(hipStreamWaitEvent(D.stream, sol.join, 0));
        }
        // info and the sum of squares of every sector's eigenvectors (NaN check), one wait
        std::vector<double> zsq(nb, 0.0);
        {
            DevBuf zs(D.pool.get(), (size_t)nb * sizeof(double)), sc(D.pool.get(), 2048 * sizeof(double));
            std::vector<double> tails(nb, 0.0);
            for (int b = 0; b < nb; ++b) {
                const int64_t cnt = (int64_t)xs[b]->k * xs[b]->k;
                const double* z = xs[b]->Zt.as<double>();
                launch_nrm2sq(D.stream, reinterpret_cast<const double2*>(z), cnt / 2, zs.as<double>() + b,
                              sc.as<double>());
                if (cnt % 2)
                    KV_HIP(hipMemcpyAsync(&tails[b], z + cnt - 1, sizeof(double), hipMemcpyDeviceToHost, D.stream));
            }
            KV_HIP(hipMemcpyAsync(zsq.data(), zs.get(), nb * sizeof(double), hipMemcpyDeviceToHost, D.stream));
            KV_HIP(hipMemcpyAsync(hinfo.data(), info.get(), nb * sizeof(rocblas_int), hipMemcpyDeviceToHost,
                                  D.stream));
            D.sync();
            for (int b = 0; b < nb; ++b) zsq[b] += tails[b] * tails[b];
        }
        if (trd_t0) {
            for (int64_t m = 1; m < N; ++m) D.trd.bytes += 16.0 * (double)(m * (m + 1) / 2) * nb;
            D.trd.columns += (long)N * nb;
        }
        for (int b = 0; b < nb; ++b) {
            const int n = xs[b]->k;
            const int64_t nn = (int64_t)n * n;
            const bool ok = hinfo[b] == 0 && std::isfinite(zsq[b]);
            if (ok) continue;
            KV_HIP(hipMemcpyAsync(W.as<double>() + (size_t)b * N, W0.as<double>() + (size_t)b * N, n * sizeof(double),
                                  hipMemcpyDeviceToDevice, D.stream));
            KV_HIP(hipMemcpyAsync(E.as<double>() + (size_t)b * N, E0.as<double>() + (size_t)b * N, n * sizeof(double),
                                  hipMemcpyDeviceToDevice, D.stream));
            const rocblas_status s2 = rocsolver_dsteqr(
                sol.h, rocblas_evect_tridiagonal, n, W.as<double>() + (size_t)b * N,
                E.as<double>() + (size_t)b * N, xs[b]->Zt.as<double>(), n, info.as<rocblas_int>() + b);
            KV_REQUIRE(s2 == rocblas_status_success, ERR_SOLVER, "rocsolver_dsteqr failed");
            read_info();
            KV_REQUIRE(hinfo[b] == 0 && std::isfinite(dev_sumsq(D, xs[b]->Zt.as<double>(), nn)), ERR_SOLVER,
                       "Hermitian eigensolver did not converge");
        }
        for (int b = 0; b < nb; ++b) {
            const int n = xs[b]->k;
            xs[b]->tau = DevBuf(D.pool.get(), (size_t)n * sizeof(double2));
            KV_HIP(hipMemcpyAsync(xs[b]->tau.get(), taub.as<double2>() + (size_t)b * N,
                                  (size_t)n * sizeof(double2), hipMemcpyDeviceToDevice, D.stream));
            // the reflectors (the sector's n x n corner), from which gram_eigvecs forms the vectors
            KV_HIP(hipMemcpy2DAsync(xs[b]->G.get(), (size_t)n * sizeof(double2), Gb.as<double2>() + b * NN,
                                    (size_t)N * sizeof(double2), (size_t)n * sizeof(double2), n,
                                    hipMemcpyDeviceToDevice, D.stream));
        }
    }
    std::vector<double> w((size_t)N * nb);
    KV_HIP(hipMemcpyAsync(w.data(), W.get(), w.size() * sizeof(double), hipMemcpyDeviceToHost, D.stream));
    D.sync();
    for (int b = 0; b < nb; ++b) {
        SectorSvd& X = *xs[b];
        const int n = X.k;
        X.s.resize(n);
        for (int j = 0; j < n; ++j) {
            const double v = w[(size_t)b * N + n - 1 - j];
            KV_REQUIRE(std::isfinite(v), ERR_SOLVER, "Hermitian eigensolver returned a non-finite value");
            X.s[j] = std::sqrt(std::max(v, 0.0) * X.gscale);
        }
        if (split) {
            X.Ev = DevBuf();
            X.ev_lo = n;  // no vectors formed yet
        } else {
            X.Ev = std::move(X.G);  // all vectors, column j = conj(v_j)
            X.ev_lo = 0;
            X.tau = DevBuf();
            X.Zt = DevBuf();
        }
    }
}

// Rows [lo, k) of a Gram sector's ascending eigenvectors into Ev: the tridiagonal
// eigenvectors Zt[:, lo:k] carried back through the reflectors, Z <- Q Z with
// Q = H_0 H_1 ... H_{n-2}, applied block by block from the last one:
// Z <- Z - V_b (T_b (V_b^H Z)) = Z - P_b Y_b,  Y_b = V_b^H Z,  P_b = V_b T_b.
// Column-major throughout (the rocBLAS layout); V_b and P_b are zero above row b*bw + 1.
// S_b = V_b^H V_b and Y_b have few outputs (bw x bw, bw x k) over up to n rows, so both run
// split over row chunks on the grouped GEMM (one group per block and chunk, partial
// products summed by a second kernel); P_b and the update Z -= P_b Y_b are rocBLAS GEMMs.
// With st and h (a second stream and the rocBLAS handle bound to it) the work goes to that
// stream, and the temporaries move into *keep, to be released once the caller has joined it.
inline void gram_eigvecs(Device& D, Solver& sol, SectorSvd& X, int lo, hipStream_t st = nullptr,
                         rocblas_handle h = nullptr, std::vector<DevBuf>* keep = nullptr) {
    if (!st) st = D.stream;
    if (!h) h = sol.h;
    const int n = X.k;
    if (X.ev_lo <= lo) return;
    KV_REQUIRE(X.Zt.get() != nullptr, ERR_INTERNAL, "Gram eigenvectors requested before the eigensolver");
    const int k = n - lo;
    DevBuf Z(D.pool.get(), (size_t)n * k * sizeof(double2));
    launch_real_cols(st, X.Zt.as<double>(), n, lo, k, n, Z.as<double2>());
    const int nref = n - 1;
    if (nref > 0 && k > 0) {
        const int bw = kReflectorBlock, kc = kReflectorBlock;  // block width, row chunk
        const int nblk = (nref + bw - 1) / bw;
        const int nch_max = (nref + kc - 1) / kc;
        const size_t vcols = (size_t)nblk * bw;
        const int64_t sbw = (int64_t)bw * bw;
        DevBuf V(D.pool.get(), (size_t)n * vcols * sizeof(double2));
        DevBuf P(D.pool.get(), (size_t)n * vcols * sizeof(double2));
        DevBuf S(D.pool.get(), (size_t)nblk * sbw * sizeof(double2));
        DevBuf T(D.pool.get(), (size_t)nblk * sbw * sizeof(double2));
        DevBuf Y(D.pool.get(), (size_t)bw * k * sizeof(double2));
        DevBuf Sp(D.pool.get(), (size_t)nch_max * nblk * sbw * sizeof(double2));
        DevBuf Yp(D.pool.get(), (size_t)nch_max * bw * k * sizeof(double2));
        launch_hh_explicit(st, X.G.as<double2>(), V.as<double2>(), n, (int)vcols);
        // work lists: S chunks of every block first, then the Y chunks block by block.
        // Row-major view: S_b^T = A B with A[c][i] = V[i, c] ('N', ld n) and B = conj(V) ('C');
        // Y_b^T = Zr conj(V_b) with Zr[j][i] = Z[i, j] ('N', ld n).
        std::vector<GemmGroup> gs;
        std::vector<GemmSeg> ss;
        std::vector<GemmTile> ts;
        auto add = [&](int m, int64_t c_off, int64_t a_off, int64_t b_off, int kk) {
            GemmGroup g{};
            g.c_off = c_off;
            g.m = m;
            g.n = bw;
            g.ldc = bw;
            g.seg_begin = (int)ss.size();
            g.seg_count = 1;
            GemmSeg sg{};
            sg.a_off = a_off;
            sg.b_off = b_off;
            sg.k = kk;
            sg.lda = n;
            sg.ldb = n;
            ss.push_back(sg);
            const int gi = (int)gs.size();
            gs.push_back(g);
            for (int a = 0; a < (m + GEMM_BM - 1) / GEMM_BM; ++a)
                for (int b = 0; b < (bw + GEMM_BN - 1) / GEMM_BN; ++b) ts.push_back(GemmTile{gi, a, b, 0});
        };
        std::vector<int> nch(nblk);
        for (int b = 0; b < nblk; ++b) {
            const int r0 = b * bw + 1, mb = n - r0;
            nch[b] = (mb + kc - 1) / kc;
            for (int c = 0; c < nch[b]; ++c)
                add(bw, ((int64_t)c * nblk + b) * sbw, (int64_t)b * bw * n + r0 + (int64_t)c * kc,
                    (int64_t)b * bw * n + r0 + (int64_t)c * kc, std::min(kc, mb - c * kc));
        }
        const int s_tiles = (int)ts.size();
        std::vector<int> y_off(nblk), y_cnt(nblk);
        for (int b = 0; b < nblk; ++b) {
            const int r0 = b * bw + 1, mb = n - r0;
            y_off[b] = (int)ts.size();
            for (int c = 0; c < nch[b]; ++c)
                add(k, (int64_t)c * bw * k, r0 + (int64_t)c * kc, (int64_t)b * bw * n + r0 + (int64_t)c * kc,
                    std::min(kc, mb - c * kc));
            y_cnt[b] = (int)ts.size() - y_off[b];
        }
        PackedUpload pk = upload_packed_on(D, st, gs, ss, ts);
        const GemmTile* tiles = pk.at<GemmTile>(2);
        launch_zero(st, Sp.get(), Sp.bytes());  // blocks with fewer chunks leave parts unset
        launch_gemm<double2>(st, 0, 2, V.as<double2>(), V.as<double2>(), Sp.as<double2>(), tiles,
                             s_tiles, pk.at<GemmGroup>(0), pk.at<GemmSeg>(1), 1.0, 0.0);
        launch_sum_parts(st, Sp.as<double2>(), nch_max, (int64_t)nblk * sbw, S.as<double2>());
        launch_hh_tfactor(st, S.as<double2>(), X.tau.as<double2>(), T.as<double2>(), nblk, bw, nref);
        auto* v = reinterpret_cast<rocblas_double_complex*>(V.get());
        auto* p = reinterpret_cast<rocblas_double_complex*>(P.get());
        auto* t = reinterpret_cast<rocblas_double_complex*>(T.get());
        auto* y = reinterpret_cast<rocblas_double_complex*>(Y.get());
        auto* z = reinterpret_cast<rocblas_double_complex*>(Z.get());
        const rocblas_double_complex one(1.0, 0.0), zero(0.0, 0.0), mone(-1.0, 0.0);
        const rocblas_stride sv = (rocblas_stride)n * bw;
        auto ok = [](rocblas_status st) {
            KV_REQUIRE(st == rocblas_status_success, ERR_SOLVER, "rocBLAS zgemm failed");
        };
        // P_b = V_b T_b for every block at once
        ok(rocblas_zgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, n, bw,
                                         bw, &one, v, n, sv, t, bw, (rocblas_stride)sbw, &zero, p, n, sv,
                                         nblk));
        for (int b = nblk - 1; b >= 0; --b) {
            const int r0 = b * bw + 1;
            const int mb = n - r0;
            if (mb <= 0) continue;
            launch_gemm<double2>(st, 0, 2, Z.as<double2>(), V.as<double2>(), Yp.as<double2>(),
                                 tiles + y_off[b], y_cnt[b], pk.at<GemmGroup>(0), pk.at<GemmSeg>(1), 1.0,
                                 0.0);
            launch_sum_parts(st, Yp.as<double2>(), nch[b], (int64_t)bw * k, Y.as<double2>());
            ok(rocblas_zgemm(h, rocblas_operation_none, rocblas_operation_none, mb, k, bw, &mone,
                             p + (size_t)b * sv + r0, n, y, bw, &one, z + r0, n));
        }
        if (keep) {
            keep->push_back(std::move(V));
            keep->push_back(std::move(P));
            keep->push_back(std::move(S));
            keep->push_back(std::move(T));
            keep->push_back(std::move(Y));
            keep->push_back(std::move(Sp));
            keep->push_back(std::move(Yp));
            keep->push_back(std::move(pk.buf));
        }
    }
    X.Ev = std::move(Z);
    X.ev_lo = lo;
}

// Form the nk leading singular vectors of a Gram-route sector.
inline void gram_vectors(Device& D, Solver& sol, SectorSvd& X, int nk) {
    const int n = X.k;
    gram_eigvecs(D, sol, X, n - nk);
    DevBuf B(D.pool.get(), (size_t)nk * n * sizeof(double2));  // rows v_j^H (or u_j^H), descending
    {
        std::vector<Copy2D> cps(nk);
        for (int j = 0; j < nk; ++j) {
            Copy2D c{};
            c.src_off = (int64_t)(n - 1 - j - X.ev_lo) * n;
            c.dst_off = (int64_t)j * n;
            c.rows = 1;
            c.cols = n;
            c.src_ld = n;
            c.dst_ld = n;
            cps[j] = c;
        }
        DevBuf dc = upload_vec(D, cps);
        launch_copy2d<double2>(D.stream, X.Ev.as<double2>(), B.as<double2>(), dc.as<Copy2D>(), nk, 1,
                               nullptr);
    }
    std::vector<double> inv(nk);
    for (int j = 0; j < nk; ++j) inv[j] = X.s[j] > 0 ? 1.0 / X.s[j] : 0.0;
    DevBuf dinv = upload_vec(D, inv);
    X.S = upload_vec(D, std::vector<double>(X.s.begin(), X.s.begin() + nk));
    // B2 = diag(1/s) B
    DevBuf B2(D.pool.get(), (size_t)nk * n * sizeof(double2));
    {
        Copy2D c{};
        c.rows = nk;
        c.cols = n;
        c.src_ld = n;
        c.dst_ld = n;
        c.scale_mode = 1;
        std::vector<Copy2D> v{c};
        DevBuf dc = upload_vec(D, v);
        launch_copy2d<double2>(D.stream, B.as<double2>(), B2.as<double2>(), dc.as<Copy2D>(), 1, nk,
                               dinv.as<double>());
    }
    X.ldU = nk;
    X.U = DevBuf(D.pool.get(), (size_t)(X.nr * nk) * sizeof(double2));
    X.Vh = DevBuf(D.pool.get(), (size_t)(nk * X.nc) * sizeof(double2));
    const double2* m = X.M.as<double2>();
    if (X.colside) {
        // V_M^H = B (nk x nc);  U_M = M (diag(1/s) B)^H  (nr x nk)
        KV_HIP(hipMemcpyAsync(X.Vh.get(), B.get(), (size_t)nk * n * sizeof(double2),
                              hipMemcpyDeviceToDevice, D.stream));
        gemm_dense(D, 'N', 'C', m, (int)X.nc, B2.as<double2>(), n, X.U.as<double2>(), nk,
                   (int)X.nr, nk, (int)X.nc);
    } else {
        // U_M = B^H (nr x nk);  V_M^H = (diag(1/s) B) M  (nk x nc)
        launch_ctranspose(D.stream, B.as<double2>(), X.U.as<double2>(), nk, n, n, nk);
        gemm_dense(D, 'N', 'N', B2.as<double2>(), n, m, (int)X.nc, X.Vh.as<double2>(), (int)X.nc,
                   nk, (int)X.nc, n);
    }
}

// Isometric route, second pass over the tail: the eigenvectors of G with values below
// tail * s_max (buffer rows [0, nt) in ascending order) span an accurate subspace; M
// restricted to it (Mt = Bt M or M Bt^H) has norm ~ tail * s_max, and the eigenvectors Y
// of its own Gram matrix rotate the tail rows: Bt <- Y^H Bt.
inline void iso_refine(Device& D, Solver& sol, SectorSvd& X, double tail, int mode, bool roc_trd) {
    X.refined = true;
    const int n = X.k;
    const int kmax = (int)std::min(X.nr, X.nc);
    const double smax = X.s.empty() ? 0.0 : X.s[0];
    int na = 0;
    while (na < kmax && X.s[na] >= tail * smax) ++na;
    const int nt = n - na;
    if (nt <= 0 || na >= kmax) return;
    gram_eigvecs(D, sol, X, 0);
    double2* Bt = X.Ev.as<double2>();  // rows [0, nt) of the eigenvector buffer (ld n)
    const double2* m = X.M.as<double2>();
    SectorSvd T;
    T.k = nt;
    T.gram = true;
    T.G = DevBuf(D.pool.get(), (size_t)nt * nt * sizeof(double2));
    // Mt is scaled to norm ~1 before its Gram matrix is formed, so the eigensolver sees the
    // same scale in both passes
    const double ref = std::max(X.s[na], tail * smax);
    const double2 inv = make_double2(ref > 0 ? 1.0 / ref : 1.0, 0.0);
    if (!X.colside) {  // rows side: Mt = Bt M (nt x nc), Gt = Mt Mt^H
        DevBuf Mt(D.pool.get(), (size_t)nt * X.nc * sizeof(double2));
        gemm_dense(D, 'N', 'N', Bt, n, m, (int)X.nc, Mt.as<double2>(), (int)X.nc, nt, (int)X.nc,
                   (int)X.nr);
        launch_zaxpby(D.stream, Mt.as<double2>(), Mt.as<double2>(), nullptr, inv,
                      make_double2(0.0, 0.0), (int64_t)nt * X.nc);
        gemm_dense(D, 'N', 'C', Mt.as<double2>(), (int)X.nc, Mt.as<double2>(), (int)X.nc,
                   T.G.as<double2>(), nt, nt, nt, (int)X.nc);
    } else {  // columns side: Mt = M Bt^H (nr x nt), Gt = Mt^H Mt
        DevBuf Mt(D.pool.get(), (size_t)X.nr * nt * sizeof(double2));
        gemm_dense(D, 'N', 'C', m, (int)X.nc, Bt, n, Mt.as<double2>(), nt, (int)X.nr, nt,
                   (int)X.nc);
        launch_zaxpby(D.stream, Mt.as<double2>(), Mt.as<double2>(), nullptr, inv,
                      make_double2(0.0, 0.0), (int64_t)X.nr * nt);
        gemm_dense(D, 'C', 'N', Mt.as<double2>(), nt, Mt.as<double2>(), nt, T.G.as<double2>(),
                   nt, nt, nt, (int)X.nr);
    }
    gram_solve(D, sol, {&T}, mode, roc_trd);
    gram_eigvecs(D, sol, T, 0);
    for (double& v : T.s) v *= (ref > 0 ? ref : 1.0);
    // Bt <- Ybuf Bt (rows y_j^H, ascending) and the tail values replaced
    DevBuf Bn(D.pool.get(), (size_t)nt * n * sizeof(double2));
    gemm_dense(D, 'N', 'N', T.Ev.as<double2>(), nt, Bt, n, Bn.as<double2>(), n, nt, n, nt);
    KV_HIP(hipMemcpyAsync(Bt, Bn.get(), (size_t)nt * n * sizeof(double2), hipMemcpyDeviceToDevice,
                          D.stream));
    for (int j = 0; j < kmax - na; ++j) X.s[na + j] = T.s[j];  // the rest is null space
}

// Isometric route: the nk leading vectors are rows of the eigenvector buffer; the other
// factor is one GEMM and carries S.
inline void iso_vectors(Device& D, Solver& sol, SectorSvd& X, int nk) {
    const int n = X.k;
    gram_eigvecs(D, sol, X, n - nk);
    DevBuf B(D.pool.get(), (size_t)nk * n * sizeof(double2));  // descending rows
    {
        std::vector<Copy2D> cps(nk);
        for (int j = 0; j < nk; ++j) {
            Copy2D c{};
            c.src_off = (int64_t)(n - 1 - j - X.ev_lo) * n;
            c.dst_off = (int64_t)j * n;
            c.rows = 1;
            c.cols = n;
            c.src_ld = n;
            c.dst_ld = n;
            cps[j] = c;
        }
        DevBuf dc = upload_vec(D, cps);
        launch_copy2d<double2>(D.stream, X.Ev.as<double2>(), B.as<double2>(), dc.as<Copy2D>(), nk, 1,
                               nullptr);
    }
    X.s_absorbed = true;
    X.S = upload_vec(D, std::vector<double>(X.s.begin(), X.s.begin() + nk));
    X.ldU = nk;
    X.U = DevBuf(D.pool.get(), (size_t)(X.nr * nk) * sizeof(double2));
    X.Vh = DevBuf(D.pool.get(), (size_t)(nk * X.nc) * sizeof(double2));
    const double2* m = X.M.as<double2>();
    if (!X.colside) {
        // U_M = B^H (nr x nk); S V^H = B M (nk x nc)
        launch_ctranspose(D.stream, B.as<double2>(), X.U.as<double2>(), nk, n, n, nk);
        gemm_dense(D, 'N', 'N', B.as<double2>(), n, m, (int)X.nc, X.Vh.as<double2>(), (int)X.nc,
                   nk, (int)X.nc, n);
    } else {
        // V^H = B (nk x nc); U S = M B^H (nr x nk)
        KV_HIP(hipMemcpyAsync(X.Vh.get(), B.get(), (size_t)nk * n * sizeof(double2),
                              hipMemcpyDeviceToDevice, D.stream));
        gemm_dense(D, 'N', 'C', m, (int)X.nc, B.as<double2>(), n, X.U.as<double2>(), nk,
                   (int)X.nr, nk, (int)X.nc);
    }
}

// ---- block subspace route -------------------------------------------------------------
// A two-site sector matrix of a D = 256 bond has sides in the thousands (2048 x 4096 for the
// 32-site chain) but keeps at most maxdim values.  Its leading singular subspace on the
// isometric side comes from a seeded random start block of b columns with power steps,
// orthonormalised by Cholesky QR between the products (randomised range finder with
// subspace iteration: Halko, Martinsson and Tropp, SIAM Review 53 (2011), algorithms
// 4.3/4.4), then Rayleigh-Ritz on the block.  b is maxdim plus kBlockOversample, trimmed to
// 256 when that still holds maxdim.  Each Ritz pair's residual |G z - theta z| (G the Gram matrix of the
// isometric side) is measured; the leading pairs within kBlockResid x the largest Ritz value
// count as singular values, and |M|_F^2 minus their sum bounds every other value of the
// sector.  A cut that keeps only certified values above the tail threshold and leaves the
// bound below it (below the cutoff, or below the smallest kept value when maxdim binds) is
// settled; otherwise the sector takes another power step, and after kBlockPowerMax of them
// the full eigensolver.  All sectors of a bond in the route run together: one grouped GEMM
// launch per product, one batched Cholesky QR, one batched Rayleigh-Ritz eigensolve.
constexpr int kBlockOversample = 32;
constexpr int kQrReduceRatio = 2;  // isometric side / other side from which qr_reduce runs
constexpr int kBlockEigSmall = 256;    // largest side of rocSOLVER's one-kernel tridiagonalisation
constexpr double kBlockDrop = 1e-12;   // Cholesky QR pivot threshold (x the largest)
constexpr double kBlockResid = 1e-13;  // Ritz residual bound (x the largest Ritz value)
constexpr int kBlockPowerMax = 3;      // power steps before the sector goes to the full route
// rounding allowance of |M|_F^2 - sum of certified theta (relative to |M|_F^2)
constexpr double kBlockMassSlack = 1e-13;

// One dense product of a batched launch: C (m x n, ld n) = op(A) op(B).
struct GemmJob {
    const double2* A;
    int lda;
    const double2* B;
    int ldb;
    double2* C;
    int m, n, k;
};

// Every job's product in one grouped GEMM launch (the jobs share opA and opB).  Operand
// offsets are measured from the lowest address among the jobs' operands, so their buffers may
// be separate allocations.  When the outputs together have too few 64 x 64 tiles to fill the
// GPU, K is cut into chunks whose partial products a second kernel sums in chunk order.
// Tile config of the block-route products, the workgroup count the K split aims for, the
// shortest K chunk it cuts and the tile order: KV_BLOCK_GEMM="cfg,wg[,kchunk[,xcd]]" (cfg an
// index of kGemmConfigs built for the op pairs CN, NN, NC; xcd 1 orders each product's tiles
// so that the tiles of one row of tiles, which read the same panel of the large operand, go
// to one XCD: workgroups are dealt to the 8 XCDs round-robin).
struct BlockGemmTuning {
    int cfg = 0, wg = 768, kchunk = 64, xcd = 0;
    static const BlockGemmTuning& get() {
        static const BlockGemmTuning t = [] {
            BlockGemmTuning o;
            if (const char* e = std::getenv("KV_BLOCK_GEMM")) {
                int c = 0, w = 0, k = 64, x = 0;
                const int got = std::sscanf(e, "%d,%d,%d,%d", &c, &w, &k, &x);
                if (got >= 2 && c >= 0 && c < kNumGemmConfigs && w > 0 && k >= 32) {
                    o.cfg = c;
                    o.wg = w;
                    o.kchunk = k;
                    o.xcd = x;
                }
            }
            return o;
        }();
        return t;
    }
};

inline void gemm_jobs(Device& D, char opA, char opB, const std::vector<GemmJob>& jobs) {
    if (jobs.empty()) return;
    const BlockGemmTuning& tune = BlockGemmTuning::get();
    const int cfg = gemm_config_supports(tune.cfg, op_code(opA), op_code(opB), true) ? tune.cfg : 0;
    const int BM = kGemmConfigs[cfg].bm, BN = kGemmConfigs[cfg].bn;
    int tiles = 0, kmin = jobs[0].k;
    const double2 *a0 = jobs[0].A, *b0 = jobs[0].B;
    double2* cmin = jobs[0].C;
    for (const auto& j : jobs) {
        KV_REQUIRE(j.k > 0 && j.m > 0 && j.n > 0, ERR_INTERNAL, "block route: empty product");
        tiles += ((j.m + BM - 1) / BM) * ((j.n + BN - 1) / BN);
        kmin = std::min(kmin, j.k);
        a0 = std::min(a0, j.A);
        b0 = std::min(b0, j.B);
        cmin = std::min(cmin, j.C);
    }
    int nch = 1;
    while (tiles * nch < tune.wg && kmin / (2 * nch) >= tune.kchunk && 2 * nch <= tune.kchunk) nch *= 2;
    size_t part_elems = 0;
    if (nch > 1)
        for (const auto& j : jobs) part_elems += (size_t)nch * j.m * j.n;
    DevBuf part(D.pool.get(), part_elems * sizeof(double2));
    double2* c0 = nch > 1 ? part.as<double2>() : cmin;
    std::vector<GemmGroup> gs;
    std::vector<GemmSeg> ss;
    std::vector<GemmTile> ts;
    std::vector<int64_t> poff(jobs.size());
    int64_t pos = 0;
    double flop = 0, bytes = 0;
    for (size_t q = 0; q < jobs.size(); ++q) {
        const GemmJob& j = jobs[q];
        // chunk depth ceil(k / nch), rounded up to a multiple of 8 where that leaves the last
        // chunk nonempty (k / nch >= kchunk >= nch keeps ceil's chunks nonempty)
        const int kq = (j.k + nch - 1) / nch, k8 = (kq + 7) / 8 * 8;
        const int kc = nch == 1 ? j.k : ((int64_t)k8 * (nch - 1) < j.k ? k8 : kq);
        poff[q] = pos;
        for (int c = 0; c < nch; ++c) {
            const int k0 = c * kc;
            GemmGroup g{};
            g.c_off = nch > 1 ? pos + (int64_t)c * j.m * j.n : (int64_t)(j.C - c0);
            g.m = j.m;
            g.n = j.n;
            g.ldc = j.n;
            g.seg_begin = (int)ss.size();
            g.seg_count = 1;
            GemmSeg sg{};
            sg.a_off = (int64_t)(j.A - a0) + (opA == 'N' ? (int64_t)k0 : (int64_t)k0 * j.lda);
            sg.b_off = (int64_t)(j.B - b0) + (opB == 'N' ? (int64_t)k0 * j.ldb : (int64_t)k0);
            sg.k = std::min(kc, j.k - k0);
            sg.lda = j.lda;
            sg.ldb = j.ldb;
            KV_REQUIRE(sg.k > 0, ERR_INTERNAL, "block route: empty K chunk");
            const int gi = (int)gs.size();
            gs.push_back(g);
            ss.push_back(sg);
            const int MT = (j.m + BM - 1) / BM, NT = (j.n + BN - 1) / BN;
            if (tune.xcd) {  // rows of tiles in groups of 8, each row's tiles 8 apart
                for (int a0 = 0; a0 < MT; a0 += 8)
                    for (int b = 0; b < NT; ++b)
                        for (int a = a0; a < std::min(a0 + 8, MT); ++a) ts.push_back(GemmTile{gi, a, b, 0});
            } else {
                for (int a = 0; a < MT; ++a)
                    for (int b = 0; b < NT; ++b) ts.push_back(GemmTile{gi, a, b, 0});
            }
        }
        if (nch > 1) pos += (int64_t)nch * j.m * j.n;
        flop += 8.0 * j.m * j.n * (double)j.k;
        bytes += 16.0 * ((double)j.m * j.k + (double)j.k * j.n + (double)j.m * j.n);
    }
    PackedUpload pk = upload_packed(D, gs, ss, ts);
    hipEvent_t t0 = D.spans.mark(D.stream);
    launch_gemm<double2>(D.stream, op_code(opA), op_code(opB), a0, b0, c0, pk.at<GemmTile>(2), (int)ts.size(),
                         pk.at<GemmGroup>(0), pk.at<GemmSeg>(1), 1.0, 0.0, cfg);
    if (nch > 1)
        for (size_t q = 0; q < jobs.size(); ++q)
            launch_sum_parts(D.stream, part.as<double2>() + poff[q], nch, (int64_t)jobs[q].m * jobs[q].n,
                             jobs[q].C);
    D.spans.close(t0, D.stream, &D.blk.gemm_ms);
    D.blk.gemm_flop += flop;
    D.blk.gemm_bytes += bytes;
    D.blk.gemm_launches += 1;
}

// One sector of a bond in the block route.  A is M (isometry on the rows) or M^H (on the
// columns), so that A's rows are the isometric side: ni x no.
struct BlockJob {
    SectorSvd* X = nullptr;
    int ni = 0, no = 0;
    uint64_t seed = 0;
    double f2 = 0;             // |M|_F^2
    DevBuf Q, P;               // isometric-side basis (ni x b) and A^H Q (no x b)
    int power = 0;             // power steps taken
    int cert = 0;              // certified Ritz pairs at the last Rayleigh-Ritz
    const char* fallback = ""; // why the sector left the route (trace output)
    std::vector<double> theta; // leading Ritz values, descending
};

// out_j = A_j in_j (herm false: ni x c) or A_j^H in_j (herm true: no x c) for every job.
inline void block_amul(Device& D, std::vector<BlockJob*>& J, bool iso_rows, bool herm,
                       const std::vector<const double2*>& in, int c, const std::vector<double2*>& out) {
    const bool conj_m = iso_rows == herm;  // the product takes M^H
    std::vector<GemmJob> g;
    for (size_t q = 0; q < J.size(); ++q) {
        const SectorSvd& X = *J[q]->X;
        const int nr = (int)X.nr, nc = (int)X.nc;
        if (conj_m)
            g.push_back(GemmJob{X.M.as<double2>(), nc, in[q], c, out[q], nc, c, nr});
        else
            g.push_back(GemmJob{X.M.as<double2>(), nc, in[q], c, out[q], nr, c, nc});
    }
    gemm_jobs(D, conj_m ? 'C' : 'N', 'N', g);
}

// Orthonormal bases of the columns of every Y_j (n_j x b, row-major), in place: Cholesky QR
// passes (one leaves an orthogonality error ~eps cond(Y)^2, a second brings it to rounding).
// Columns whose pivot falls to kBlockDrop x the largest depend on the earlier ones: with
// `refill` they are replaced by seeded random vectors at the scale of the largest column and
// the pass repeats (at most twice), so the basis keeps b independent columns; without, they
// come out zero and stay zero through the products that follow.
inline void block_orth(Device& D, std::vector<DevBuf*>& Y, const std::vector<int>& n, int b, int passes,
                       bool refill, uint64_t seed) {
    const int nj = (int)Y.size();
    KV_REQUIRE(b <= kCholMax, ERR_INTERNAL, "block route: block wider than kCholMax");
    // S_j = Y_j^H Y_j (one grouped GEMM), the left-looking panel Cholesky with T_j = R_j^{-1}
    // (two launches per 32 columns, all jobs together), Q_j = Y_j T_j (one grouped GEMM)
    const size_t bb = (size_t)b * b;
    DevBuf S(D.pool.get(), nj * bb * sizeof(double2)), T(D.pool.get(), nj * bb * sizeof(double2));
    DevBuf Rd(D.pool.get(), (size_t)chol_left_rd_elems(b, nj) * sizeof(double2));
    DevBuf live(D.pool.get(), (size_t)nj * b * sizeof(int)), dmax(D.pool.get(), (size_t)nj * sizeof(double));
    DevBuf sz = upload_vec(D, std::vector<int>(nj, b));
    std::vector<int> hl((size_t)nj * b);
    for (int pass = 0; pass < passes; ++pass) {
        for (int attempt = 0;; ++attempt) {
            std::vector<GemmJob> g;
            for (int q = 0; q < nj; ++q)
                g.push_back(GemmJob{Y[q]->as<double2>(), b, Y[q]->as<double2>(), b, S.as<double2>() + q * bb, b, b,
                                    n[q]});
            gemm_jobs(D, 'C', 'N', g);
            hipEvent_t t0 = D.spans.mark(D.stream);
            launch_chol_left(D.stream, S.as<double2>(), sz.as<int>(), b, b, (int64_t)bb, kBlockDrop, live.as<int>(),
                             dmax.as<double>(), T.as<double2>(), Rd.as<double2>(), nj);
            D.spans.close(t0, D.stream, &D.blk.orth_ms);
            D.blk.orth_calls += 1;
            // dependent columns stay zero: the host needs no flags.  A second pass factors the
            // Gram matrix of an orthonormal basis (pivots near one), so only the first checks.
            if (!refill || pass > 0) break;
            KV_HIP(hipMemcpyAsync(hl.data(), live.get(), hl.size() * sizeof(int), hipMemcpyDeviceToHost, D.stream));
            D.sync();
            bool dead = false;
            for (int v : hl) dead = dead || v == 0;
            if (!dead || attempt == 2) break;
            for (int q = 0; q < nj; ++q)
                launch_refill_cols(D.stream, Y[q]->as<double2>(), n[q], b, live.as<int>() + (size_t)q * b,
                                   dmax.as<double>() + q, seed + 977u * (uint64_t)(pass * 3 + attempt + 1) + q);
        }
        std::vector<DevBuf> Qn;
        std::vector<GemmJob> g;
        for (int q = 0; q < nj; ++q) {
            Qn.emplace_back(D.pool.get(), (size_t)n[q] * b * sizeof(double2));
            g.push_back(GemmJob{Y[q]->as<double2>(), b, T.as<double2>() + q * bb, b, Qn.back().as<double2>(), n[q], b,
                                b});
        }
        gemm_jobs(D, 'N', 'N', g);
        for (int q = 0; q < nj; ++q) *Y[q] = std::move(Qn[q]);
    }
}

// Start blocks: Q_j = orth(A_j Omega_j) with Omega_j (no x b) seeded uniform, one pass.
inline void block_start(Device& D, std::vector<BlockJob*>& J, bool iso_rows, int b) {
    std::vector<DevBuf> Om;
    std::vector<const double2*> in;
    std::vector<double2*> out;
    std::vector<DevBuf*> ys;
    std::vector<int> n;
    for (auto* j : J) {
        Om.emplace_back(D.pool.get(), (size_t)j->no * b * sizeof(double2));
        launch_rand_fill(D.stream, Om.back().as<double2>(), (int64_t)j->no * b, j->seed);
        j->Q = DevBuf(D.pool.get(), (size_t)j->ni * b * sizeof(double2));
    }
    for (size_t q = 0; q < J.size(); ++q) {
        in.push_back(Om[q].as<double2>());
        out.push_back(J[q]->Q.as<double2>());
        ys.push_back(&J[q]->Q);
        n.push_back(J[q]->ni);
    }
    block_amul(D, J, iso_rows, false, in, b, out);
    block_orth(D, ys, n, b, 1, false, J[0]->seed ^ 0x5bd1e995u);
}

// One power step for every job, Q <- orth(A orth(A^H Q)), then P = A^H Q for Rayleigh-Ritz.
inline void block_power(Device& D, std::vector<BlockJob*>& J, bool iso_rows, int b) {
    std::vector<const double2*> in;
    std::vector<double2*> out;
    std::vector<DevBuf*> ys;
    std::vector<int> n;
    for (auto* j : J) {
        j->P = DevBuf(D.pool.get(), (size_t)j->no * b * sizeof(double2));
        in.push_back(j->Q.as<double2>());
        out.push_back(j->P.as<double2>());
        ys.push_back(&j->P);
        n.push_back(j->no);
    }
    block_amul(D, J, iso_rows, true, in, b, out);
    block_orth(D, ys, n, b, 1, false, J[0]->seed ^ 0x27d4eb2fu);
    in.clear();
    out.clear();
    ys.clear();
    n.clear();
    for (auto* j : J) {
        j->Q = DevBuf(D.pool.get(), (size_t)j->ni * b * sizeof(double2));
        in.push_back(j->P.as<double2>());
        out.push_back(j->Q.as<double2>());
        ys.push_back(&j->Q);
        n.push_back(j->ni);
    }
    block_amul(D, J, iso_rows, false, in, b, out);
    block_orth(D, ys, n, b, 2, true, J[0]->seed ^ (0x9e3779b9u + 31u * (uint64_t)J[0]->power));
    in.clear();
    out.clear();
    for (auto* j : J) {
        j->P = DevBuf(D.pool.get(), (size_t)j->no * b * sizeof(double2));
        in.push_back(j->Q.as<double2>());
        out.push_back(j->P.as<double2>());
        j->power += 1;
    }
    block_amul(D, J, iso_rows, true, in, b, out);
}

// Rayleigh-Ritz on every job's block: H = P^H P / |M|_F^2 = V Theta V^H (one batched
// eigensolve), Z = Q V and W = P V for all b pairs, their residuals, the certified values
// (X.s) and the bound on every other value (X.ceil).
inline void block_ritz(Device& D, Solver& sol, std::vector<BlockJob*>& J, bool iso_rows, int b, int mode,
                       bool roc_trd) {
    const int nj = (int)J.size();
    const int kc = b;
    std::vector<SectorSvd> H(nj);
    std::vector<SectorSvd*> hp;
    std::vector<GemmJob> g;
    for (int q = 0; q < nj; ++q) {
        H[q].k = b;
        H[q].gram = true;
        H[q].G = DevBuf(D.pool.get(), (size_t)b * b * sizeof(double2));
        g.push_back(GemmJob{J[q]->P.as<double2>(), b, J[q]->P.as<double2>(), b, H[q].G.as<double2>(), b, b, J[q]->no});
        hp.push_back(&H[q]);
    }
    gemm_jobs(D, 'C', 'N', g);
    for (int q = 0; q < nj; ++q)
        launch_zaxpby(D.stream, H[q].G.as<double2>(), H[q].G.as<double2>(), nullptr,
                      make_double2(1.0 / J[q]->f2, 0.0), make_double2(0.0, 0.0), (int64_t)b * b);
    {
        hipEvent_t t0 = D.spans.mark(D.stream);
        // the split route (one-workgroup tridiagonalisation, dstedc, back-transformation):
        // rocSOLVER's zheevd spends most of its time reducing a 256-wide matrix column by column
        gram_solve(D, sol, hp, mode == EIG_HEEVD ? EIG_HEEVD : EIG_SPLIT, roc_trd, true);
        if (hp.size() == 2 && mode != EIG_HEEVD) {
            // the pair's back-transformations side by side: the second sector first, on the
            // second stream (its temporaries are allocated before the first sector's and kept
            // until the join), then the first on the context stream
            std::vector<DevBuf> keep;
            sol.second();
            KV_HIP(hipEventRecord(sol.fork, D.stream));
            KV_HIP(hipStreamWaitEvent(sol.s2, sol.fork, 0));
            gram_eigvecs(D, sol, *hp[1], 0, sol.s2, sol.h2, &keep);
            gram_eigvecs(D, sol, *hp[0], 0);
            KV_HIP(hipEventRecord(sol.join, sol.s2));
            KV_HIP(hipStreamWaitEvent(D.stream, sol.join, 0));
        } else {
            for (auto* h : hp) gram_eigvecs(D, sol, *h, 0);
        }
        D.spans.close(t0, D.stream, &D.blk.eig_ms);
        D.blk.eig_calls += 1;
    }
    std::vector<DevBuf> Vd;
    for (int q = 0; q < nj; ++q) {  // rows v_j^H in descending order
        Vd.emplace_back(D.pool.get(), (size_t)kc * b * sizeof(double2));
        std::vector<Copy2D> cps(kc);
        for (int j = 0; j < kc; ++j) {
            Copy2D c{};
            c.src_off = (int64_t)(b - 1 - j - H[q].ev_lo) * b;
            c.dst_off = (int64_t)j * b;
            c.rows = 1;
            c.cols = b;
            c.src_ld = b;
            c.dst_ld = b;
            cps[j] = c;
        }
        DevBuf dc = upload_vec(D, cps);
        launch_copy2d<double2>(D.stream, H[q].Ev.as<double2>(), Vd.back().as<double2>(), dc.as<Copy2D>(), kc, 1,
                               nullptr);
    }
    std::vector<GemmJob> gz, gw;
    for (int q = 0; q < nj; ++q) {
        SectorSvd& X = *J[q]->X;
        X.Z = DevBuf(D.pool.get(), (size_t)J[q]->ni * kc * sizeof(double2));
        X.W = DevBuf(D.pool.get(), (size_t)J[q]->no * kc * sizeof(double2));
        gz.push_back(GemmJob{J[q]->Q.as<double2>(), b, Vd[q].as<double2>(), b, X.Z.as<double2>(), J[q]->ni, kc, b});
        gw.push_back(GemmJob{J[q]->P.as<double2>(), b, Vd[q].as<double2>(), b, X.W.as<double2>(), J[q]->no, kc, b});
    }
    gemm_jobs(D, 'N', 'C', gz);
    gemm_jobs(D, 'N', 'C', gw);
    // residuals of the Ritz pairs: A W_j = G z_j against theta_j z_j
    std::vector<DevBuf> GZ;
    std::vector<const double2*> in;
    std::vector<double2*> out;
    for (int q = 0; q < nj; ++q) {
        GZ.emplace_back(D.pool.get(), (size_t)J[q]->ni * kc * sizeof(double2));
        in.push_back(J[q]->X->W.as<double2>());
        out.push_back(GZ.back().as<double2>());
    }
    block_amul(D, J, iso_rows, false, in, kc, out);
    constexpr int rows_per = 256;
    std::vector<int> chunks(nj);
    std::vector<size_t> off(nj + 1, 0);
    std::vector<std::vector<double>> theta(nj, std::vector<double>(kc));
    std::vector<double> th_all;
    for (int q = 0; q < nj; ++q) {
        chunks[q] = (J[q]->ni + rows_per - 1) / rows_per;
        off[q + 1] = off[q] + (size_t)chunks[q] * kc;
        for (int j = 0; j < kc; ++j) theta[q][j] = H[q].s[j] * H[q].s[j] * J[q]->f2;
        th_all.insert(th_all.end(), theta[q].begin(), theta[q].end());
    }
    DevBuf dth = upload_vec(D, th_all);
    DevBuf part(D.pool.get(), off[nj] * sizeof(double));
    for (int q = 0; q < nj; ++q)
        launch_resid_cols(D.stream, GZ[q].as<double2>(), kc, J[q]->X->Z.as<double2>(), kc,
                          dth.as<double>() + (size_t)q * kc, J[q]->ni, kc, rows_per, part.as<double>() + off[q]);
    std::vector<double> hpart(off[nj]);
    KV_HIP(hipMemcpyAsync(hpart.data(), part.get(), hpart.size() * sizeof(double), hipMemcpyDeviceToHost, D.stream));
    D.sync();
    for (int q = 0; q < nj; ++q) {
        SectorSvd& X = *J[q]->X;
        const double bound = kBlockResid * theta[q][0];
        int cert = 0;
        for (; cert < kc; ++cert) {
            double r2s = 0.0;
            for (int y = 0; y < chunks[q]; ++y) r2s += hpart[off[q] + (size_t)y * kc + cert];
            if (!(std::sqrt(r2s) <= bound)) break;
        }
        // the sector's Frobenius mass outside the certified pairs bounds every other value
        double mass = J[q]->f2;
        for (int j = 0; j < cert; ++j) mass -= theta[q][j];
        X.ceil = std::sqrt(std::max(mass, 0.0) + kBlockMassSlack * J[q]->f2);
        X.s.assign((size_t)cert, 0.0);
        for (int j = 0; j < cert; ++j) X.s[j] = std::sqrt(theta[q][j]);
        X.block = cert > 0;
        X.iso = true;
        X.gram = false;
        X.colside = !iso_rows;
        X.kc = kc;
        X.k = cert;
        J[q]->cert = cert;
        J[q]->theta = theta[q];
    }
}

// Factors of a block-route sector: the isometry is Z's leading columns, the other factor W
// (which already carries S), conjugate-transposed where the layout needs rows.
inline void block_vectors(Device& D, SectorSvd& X, int nk) {
    const int kc = X.kc;
    X.s_absorbed = true;
    X.S = upload_vec(D, std::vector<double>(X.s.begin(), X.s.begin() + nk));
    X.ldU = kc;
    X.Vh = DevBuf(D.pool.get(), (size_t)kc * X.nc * sizeof(double2));
    if (!X.colside) {  // U = Z (nr x kc); S V^H = W^H (kc x nc)
        launch_ctranspose(D.stream, X.W.as<double2>(), X.Vh.as<double2>(), (int)X.nc, kc, kc, (int)X.nc);
        X.U = std::move(X.Z);
    } else {  // V^H = Z^H (kc x nc); U S = M V = W (nr x kc)
        launch_ctranspose(D.stream, X.Z.as<double2>(), X.Vh.as<double2>(), (int)X.nc, kc, kc, (int)X.nc);
        X.U = std::move(X.W);
    }
}

inline SvdResult svd_truncate(const std::shared_ptr<Device>& dev, Solver& sol, const Tensor& A,
                              int r1, double cutoff, int maxdim, bool left2right,
                              int left_dir = -1, const SvdOptions& opt = SvdOptions::from_env()) {
    Device& D = *dev;
    const SvdStructure S = svd_structure(A.lay, r1);
    const int np = (int)S.pairs.size();
    std::vector<SectorSvd> X(np);
    SvdResult res;
    // full routes: the isometric-side Gram matrix when the isometry is the smaller side, the
    // smaller side's Gram matrix (guarded below) otherwise, the exact SVD for small sectors
    auto full_route = [&](int p) {
        const int64_t iso_side = left2right ? X[p].nr : X[p].nc;
        const int64_t oth_side = left2right ? X[p].nc : X[p].nr;
        const bool big = !opt.force_exact && std::min(X[p].nr, X[p].nc) >= opt.iso_min;
        if (big && iso_side <= oth_side)
            gram_prepare(D, X[p], true, !left2right);
        else if (big && cutoff >= 0.0)
            gram_prepare(D, X[p]);
        else
            exact_sector(D, sol, X[p], cutoff < 0.0);
    };
    // the isometric Gram route on the longer side, Householder-reduced to the shorter one
    // when it is at least twice as long
    auto iso_route = [&](int p) {
        const int64_t iso_side = left2right ? X[p].nr : X[p].nc;
        const int64_t oth_side = left2right ? X[p].nc : X[p].nr;
        if (opt.qr_reduce && iso_side >= kQrReduceRatio * oth_side) qr_reduce(D, sol, X[p], !left2right);
        gram_prepare(D, X[p], true, !left2right);
    };
    // the block route's width: maxdim plus the oversampling, trimmed to kBlockEigSmall when
    // that still holds maxdim (rocSOLVER's zheevd tridiagonalises sides up to 256 in one
    // kernel and goes column by column above)
    const int bw = std::max(maxdim, std::min(maxdim + kBlockOversample, kBlockEigSmall));
    const bool block_ok = opt.block && !opt.force_exact && cutoff >= 0.0 && maxdim > 0 && bw <= kCholMax;
    std::vector<int> cand;
    // phase 1: gather every sector matrix and compute its singular values
    for (int p = 0; p < np; ++p) {
        const SvdSector& rs = S.rows[S.pairs[p].first];
        const SvdSector& cs = S.cols[S.pairs[p].second];
        X[p].nr = rs.len;
        X[p].nc = cs.len;
        KV_REQUIRE(rs.len < (1 << 30) && cs.len < (1 << 30), ERR_ARG, "sector matrix too large");
        X[p].M = DevBuf(D.pool.get(), (size_t)(rs.len * cs.len) * sizeof(double2));
        launch_zero(D.stream, X[p].M.get(), X[p].M.bytes());
        std::vector<Copy2D> cps;
        int maxrows = 1;
        for (const auto& rsub : rs.subs)
            for (const auto& csub : cs.subs) {
                Idx q{};
                for (int l = 0; l < S.r1; ++l) q[l] = rsub.qn[l];
                for (int l = 0; l < S.r2; ++l) q[S.r1 + l] = csub.qn[l];
                const int bi = A.lay.find(q);
                if (bi < 0) continue;
                Copy2D c{};
                c.src_off = A.lay.blocks[bi].off;
                c.dst_off = rsub.pos * cs.len + csub.pos;
                c.rows = (int)rsub.size;
                c.cols = (int)csub.size;
                c.src_ld = (int)csub.size;
                c.dst_ld = (int)cs.len;
                maxrows = std::max(maxrows, c.rows);
                cps.push_back(c);
            }
        if (!cps.empty()) {
            DevBuf dc = upload_vec(D, cps);
            launch_copy2d<double2>(D.stream, A.d(), X[p].M.as<double2>(), dc.as<Copy2D>(),
                                   (int)cps.size(), maxrows, nullptr);
        }
        // sectors at least block_min (3) times as wide as the block take the block subspace route
        if (block_ok && (double)std::min(rs.len, cs.len) >= opt.block_min * bw)
            cand.push_back(p);
        else
            full_route(p);
    }
    // the block route, all its sectors together: start block, one power step, Rayleigh-Ritz
    std::vector<BlockJob> jobs(cand.size());
    std::vector<int> job_of(np, -1);
    if (!cand.empty()) {
        std::vector<BlockJob*> jp;
        DevBuf sc(D.pool.get(), 1024 * sizeof(double)), f2d(D.pool.get(), cand.size() * sizeof(double));
        for (size_t i = 0; i < cand.size(); ++i) {
            const int p = cand[i];
            BlockJob& j = jobs[i];
            j.X = &X[p];
            j.ni = (int)(left2right ? X[p].nr : X[p].nc);
            j.no = (int)(left2right ? X[p].nc : X[p].nr);
            j.seed = ((uint64_t)X[p].nr << 40) ^ ((uint64_t)X[p].nc << 16) ^ (uint64_t)p;
            launch_nrm2sq(D.stream, X[p].M.as<double2>(), X[p].nr * X[p].nc, f2d.as<double>() + i, sc.as<double>());
            job_of[p] = (int)i;
            jp.push_back(&j);
        }
        std::vector<double> f2(cand.size());
        KV_HIP(hipMemcpyAsync(f2.data(), f2d.get(), f2.size() * sizeof(double), hipMemcpyDeviceToHost, D.stream));
        D.sync();
        for (size_t i = 0; i < cand.size(); ++i) jobs[i].f2 = f2[i] > 0 ? f2[i] : 1.0;
        block_start(D, jp, left2right, bw);
        block_power(D, jp, left2right, bw);
        block_ritz(D, sol, jp, left2right, bw, opt.eig_mode, opt.roc_trd);
        for (int p : cand)
            if (!X[p].block) {  // no certified value: the full route
                jobs[job_of[p]].fallback = " fallback: no certified value";
                X[p].Z.reset();
                X[p].W.reset();
                full_route(p);
                res.block_fallbacks += 1;
                D.blk.fallbacks += 1;
            }
    }
    // Gram eigenproblems: one call for all sectors (zheevd needs one side per call)
    {
        std::map<int, std::vector<SectorSvd*>> by_n;
        std::vector<SectorSvd*> all;
        for (int p = 0; p < np; ++p)
            if (X[p].gram) {
                by_n[X[p].k].push_back(&X[p]);
                all.push_back(&X[p]);
            }
        if (opt.eig_mode == EIG_HEEVD || by_n.size() <= 1)
            for (auto& kvp : by_n) gram_solve(D, sol, kvp.second, opt.eig_mode, opt.roc_trd);
        else
            gram_solve(D, sol, all, opt.eig_mode, opt.roc_trd);
        // the isometric route's values: only min(rows, cols) are singular values
        for (int p = 0; p < np; ++p)
            if (X[p].iso && !X[p].block) X[p].s.resize((size_t)std::min(X[p].nr, X[p].nc));
    }
    // phase 2: global truncation; Gram sectors that would keep values too small for the
    // Gram eigenvalues' absolute accuracy move to the exact route and the rule reruns
    std::vector<int> kept;
    // every redo moves one sector forward through block -> full isometric route -> refined
    // tail (or smaller-side Gram -> isometric), so the rule settles
    for (int pass = 0;; ++pass) {
        KV_REQUIRE(pass <= 4 * np + 4, ERR_INTERNAL, "bond truncation did not settle");
        std::vector<std::vector<double>> sv(np);
        for (int p = 0; p < np; ++p) sv[p] = X[p].s;
        kept = truncation_counts(sv, cutoff, maxdim);
        // smallest kept value overall: an isometric-route tail that may reach it is refined
        double s_last = -1.0;
        int tot = 0;
        for (int p = 0; p < np; ++p) {
            tot += kept[p];
            if (kept[p] > 0 && (s_last < 0 || X[p].s[kept[p] - 1] < s_last)) s_last = X[p].s[kept[p] - 1];
        }
        bool redo = false;
        std::vector<BlockJob*> more;  // block sectors that take another power step
        for (int p = 0; p < np; ++p) {
            if (X[p].block) {
                // settled: the kept values stay above the tail, and every value beyond the
                // certified ones (at most ceil) falls below the cutoff or, when maxdim binds,
                // below the smallest kept value
                const bool tail_ok = !(kept[p] > 0 && X[p].s[kept[p] - 1] < opt.tail * X[p].s[0]);
                const bool settled = X[p].ceil <= cutoff || (tot >= maxdim && X[p].ceil < s_last);
                if (tail_ok && settled) continue;
                // another power step can only tighten the bound of a maxdim-bound cut
                BlockJob& j = jobs[job_of[p]];
                if (tail_ok && tot >= maxdim && j.power < kBlockPowerMax) {
                    more.push_back(&j);
                    continue;
                }
                // the full eigensolver on the isometric side
                jobs[job_of[p]].fallback = tail_ok ? " fallback: power steps exhausted" : " fallback: tail";
                X[p].block = false;
                X[p].Z.reset();
                X[p].W.reset();
                iso_route(p);
                gram_solve(D, sol, {&X[p]}, opt.eig_mode, opt.roc_trd);
                X[p].s.resize((size_t)std::min(X[p].nr, X[p].nc));
                res.block_fallbacks += 1;
                D.blk.fallbacks += 1;
                redo = true;
            } else if (X[p].iso) {
                if (cutoff >= 0.0 && !X[p].refined && !X[p].s.empty() && s_last >= 0.0 &&
                    opt.tail * X[p].s[0] >= s_last) {
                    iso_refine(D, sol, X[p], opt.tail, opt.eig_mode, opt.roc_trd);
                    redo = true;
                }
            } else if (X[p].gram && kept[p] > 0 && X[p].s[kept[p] - 1] < opt.guard * X[p].s[0]) {
                // U = M V / s would lose orthonormality at these values: redo the sector
                // on its isometric (larger) side
                iso_route(p);
                gram_solve(D, sol, {&X[p]}, opt.eig_mode, opt.roc_trd);
                X[p].s.resize((size_t)std::min(X[p].nr, X[p].nc));
                redo = true;
            }
        }
        if (!more.empty()) {
            block_power(D, more, left2right, bw);
            block_ritz(D, sol, more, left2right, bw, opt.eig_mode, opt.roc_trd);
            for (auto* j : more)
                if (!j->X->block) {  // lost every certified value: the full route
                    const int p = (int)(j->X - X.data());
                    j->fallback = " fallback: no certified value after a power step";
                    X[p].Z.reset();
                    X[p].W.reset();
                    full_route(p);
                    if (X[p].gram) gram_solve(D, sol, {&X[p]}, opt.eig_mode, opt.roc_trd);
                    if (X[p].iso) X[p].s.resize((size_t)std::min(X[p].nr, X[p].nc));
                    res.block_fallbacks += 1;
                    D.blk.fallbacks += 1;
                }
            redo = true;
        }
        if (!redo) break;
    }
    for (int p = 0; p < np; ++p) res.kept_by_q[S.rows[S.pairs[p].first].q] = kept[p];
    for (int p = 0; p < np; ++p) {
        res.total += (int)X[p].s.size();
        res.refined_sectors += X[p].refined ? 1 : 0;
        char buf[160];
        std::snprintf(buf, sizeof buf, " [%lldx%lld %s%s smax %.2e kept %d last %.2e]",
                      (long long)S.rows[S.pairs[p].first].len, (long long)S.cols[S.pairs[p].second].len,
                      X[p].block ? "block" : X[p].iso ? (X[p].refined ? "iso2" : "iso") : (X[p].gram ? "gram" : "exact"),
                      X[p].reduced ? "-qr" : "",
                      X[p].s.empty() ? 0.0 : X[p].s[0], kept[p],
                      kept[p] > 0 ? X[p].s[kept[p] - 1] : 0.0);
        res.detail += buf;
        if (job_of[p] >= 0) {
            const BlockJob& j = jobs[job_of[p]];
            std::snprintf(buf, sizeof buf, "{power %d cert %d ceil %.2e%s}", j.power, j.cert,
                          X[p].block ? X[p].ceil : 0.0, j.fallback);
            res.detail += buf;
        }
        if (X[p].block) {
            res.block_sectors += 1;
            D.blk.sectors += 1;
            if (kept[p] > 0) block_vectors(D, X[p], kept[p]);
        } else if (X[p].iso) {
            res.gram_sectors += 1;
            if (kept[p] > 0) iso_vectors(D, sol, X[p], kept[p]);
        } else if (X[p].gram) {
            res.gram_sectors += 1;
            if (kept[p] > 0) gram_vectors(D, sol, X[p], kept[p]);
        } else {
            res.exact_sectors += 1;
        }
        if (X[p].reduced && kept[p] > 0) qr_expand(D, X[p], X[p].colside, kept[p]);
    }
    // phase 3: output layouts and scatter with S absorbed
    Layout Ll, Rl;
    Ll.rank = S.r1 + 1;
    Rl.rank = S.r2 + 1;
    for (int l = 0; l < S.r1; ++l) Ll.dirs[l] = A.lay.dirs[l];
    Ll.dirs[S.r1] = left_dir;
    Rl.dirs[0] = -left_dir;
    for (int l = 0; l < S.r2; ++l) Rl.dirs[1 + l] = A.lay.dirs[S.r1 + l];
    for (int p = 0; p < np; ++p) {
        if (kept[p] <= 0) continue;
        const SvdSector& rs = S.rows[S.pairs[p].first];
        const SvdSector& cs = S.cols[S.pairs[p].second];
        for (const auto& rsub : rs.subs) {
            BlockDesc b;
            for (int l = 0; l < S.r1; ++l) {
                b.qn[l] = rsub.qn[l];
                b.shape[l] = rsub.shape[l];
            }
            b.qn[S.r1] = -left_dir * rs.q;
            b.shape[S.r1] = kept[p];
            Ll.blocks.push_back(b);
        }
        for (const auto& csub : cs.subs) {
            BlockDesc b;
            b.qn[0] = left_dir * cs.q;
            b.shape[0] = kept[p];
            for (int l = 0; l < S.r2; ++l) {
                b.qn[1 + l] = csub.qn[l];
                b.shape[1 + l] = csub.shape[l];
            }
            Rl.blocks.push_back(b);
        }
        res.kept += kept[p];
    }
    Ll.finalize();
    Rl.finalize();
    res.left = Tensor::create(dev, Ll, false);
    res.right = Tensor::create(dev, Rl, false);
    res.svals.resize(np);
    for (int p = 0; p < np; ++p) {
        if (kept[p] <= 0) continue;
        const SvdSector& rs = S.rows[S.pairs[p].first];
        const SvdSector& cs = S.cols[S.pairs[p].second];
        const int nk = kept[p];
        const int64_t nc = cs.len;
        std::vector<Copy2D> lc, rc;
        int maxr_l = 1, maxr_r = 1;
        for (const auto& rsub : rs.subs) {
            Idx q{};
            for (int l = 0; l < S.r1; ++l) q[l] = rsub.qn[l];
            q[S.r1] = -left_dir * rs.q;
            const int bi = Ll.find(q);
            Copy2D c{};
            c.src_off = rsub.pos * (int64_t)X[p].ldU;  // U_M rows [pos, pos+size)
            c.dst_off = Ll.blocks[bi].off;
            c.rows = (int)rsub.size;
            c.cols = nk;
            c.src_ld = X[p].ldU;
            c.dst_ld = nk;
            c.scale_mode = (left2right || X[p].s_absorbed) ? 0 : 2;  // U S when S goes left
            maxr_l = std::max(maxr_l, c.rows);
            lc.push_back(c);
        }
        for (const auto& csub : cs.subs) {
            Idx q{};
            q[0] = left_dir * cs.q;
            for (int l = 0; l < S.r2; ++l) q[1 + l] = csub.qn[l];
            const int bi = Rl.find(q);
            Copy2D c{};
            c.src_off = csub.pos;  // V_M^H rows [0, nk), cols [pos, pos+size), ld nc
            c.dst_off = Rl.blocks[bi].off;
            c.rows = nk;
            c.cols = (int)csub.size;
            c.src_ld = (int)nc;
            c.dst_ld = (int)csub.size;
            c.scale_mode = (left2right && !X[p].s_absorbed) ? 1 : 0;  // S V^H when S goes right
            maxr_r = std::max(maxr_r, c.rows);
            rc.push_back(c);
        }
        PackedUpload pk = upload_packed(D, lc, rc);
        launch_copy2d<double2>(D.stream, X[p].U.as<double2>(), res.left->d(), pk.at<Copy2D>(0),
                               (int)lc.size(), maxr_l, X[p].S.as<double>());
        launch_copy2d<double2>(D.stream, X[p].Vh.as<double2>(), res.right->d(), pk.at<Copy2D>(1),
                               (int)rc.size(), maxr_r, X[p].S.as<double>());
        res.svals[p].assign(X[p].s.begin(), X[p].s.begin() + nk);
    }
    return res;
}

}  // namespace kv
