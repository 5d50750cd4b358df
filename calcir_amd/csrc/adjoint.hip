// adjoint.hip -- gradient of the IR lineshape with respect to the map
// coefficients (gfx950).
//
// The IR correlation function is C(t_k) = sum_s sum_a mu_a(t_k)^T F_s(t_k)
// mu_a(t_s), with F_s(t_k) = U(t_k) ... U(t_s+1), U(t) = exp(+i H(t) dt/hbar)
// (right endpoint, calcIR.cu:577-617, 628-648), decayed and scaled by
// 1/nsamples (calcIR.cu:777-781) and sent through the C2R of
// calcIR.cu:788-827.  H(t) and mu(t) depend on the map coefficients
// (calcIR.cu:1203-1404); the field E does not.
//
// Forward: every distinct frame is diagonalised once (batched) and its
// eigensystem kept in HBM; the dipole vectors f_{s,k} = F_s(t_k) mu(t_s)
// (3 complex columns per sample and point) are propagated frame by frame in
// the vector form and kept too.
//
// Backward (reverse sweep over the frames): with w_k the complex weight
// dL/dRe C_k - i dL/dIm C_k, each sample carries b = sum_{k'>=k} w_k'
// U(t_k+1)^T...U(t_k')^T mu(t_k'); since U is complex symmetric the sweep
// is b <- U(t) b + w_k mu(t).  At frame t the derivative of
// Re b^T U f with respect to H is, by the Daleckii-Krein formula,
//   G = V Re(D o (V^T b)(V^T f)^T) V^T,
//   D_kl = i tau e^{i tau (l_k + l_l)/2} sinc(tau (l_k - l_l)/2),
// which is finite for degenerate eigenvalues.  The per-frame products run on
// the batched GEMMs (MFMA in FP32); G is then contracted with dH/dcoef on the
// fly, recomputing the transition-dipole geometry, and the dipole terms
// (mu = u x mu') are added per chromophore.

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "ctx.hpp"
#include "gemm.hpp"
#include "sites.hpp"
#include "spectra.hpp"

namespace cir {

constexpr int kAdjNT = 256;
constexpr int kCoefRows = 16;  // rows of G per workgroup in the contraction

struct AdjState {
    Schedule S;
    int nu = 0, mmax = 1;
    // store: eigensystems of every distinct frame kept (V, w: nu frames);
    // else V, w hold one chunk of B frames, recomputed in the reverse sweep
    bool store = true;
    int B = 1;
    DevBuf V, w, sites, box;
    DevBuf H, work;  // chunk Hamiltonians and eigensolver workspace
    // f_{s,k}: ns*nt blocks of N x 6 reals (Re a=0..2, Im a=0..2 per row)
    DevBuf fv;
    // pair lists (device) and per-frame offsets (host)
    DevBuf pairs;
    std::vector<int> off_all, off_prop;  // nu+1 each
    std::vector<double> tcf;             // 2*nt: summed over samples, before decay
    bool have = false;
};

void adj_release(AdjState* a) { delete a; }

namespace {

// Work of propagate() for any 1..m column groups of 3 (the split count
// depends on the width, so the size is not monotone in it).
size_t prop_work_max(int N, int m) {
    size_t w = 0;
    for (int q = 1; q <= m; ++q) w = std::max(w, propagate_work_elems(N, 3 * q));
    return w;
}

// pair list layout in AdjState::pairs (ints): all pairs of a frame as
// (s, k, block) triples from 3*off_all[u]; propagated pairs (k >= 1) as
// (s, block(k-1), block(k)) triples from 3*(P0 + off_prop[u]).
struct PairLists {
    std::vector<int> data;
    std::vector<int> off_all, off_prop;
    int prop_base = 0, mmax = 1;
};

PairLists build_pairs(const cir_params& p, const Schedule& S) {
    const int ns = p.nsamples, nt = p.ntcfpoints, nu = int(S.uniq.size());
    std::vector<std::vector<int>> at(nu);
    for (int s = 0; s < ns; ++s)
        for (int k = 0; k < nt; ++k) at[S.uof[S.fidx[size_t(s) * nt + k]]].push_back(s * nt + k);
    PairLists L;
    L.off_all.assign(nu + 1, 0);
    L.off_prop.assign(nu + 1, 0);
    std::vector<int> all, prop;
    for (int u = 0; u < nu; ++u) {
        for (int sk : at[u]) {
            const int s = sk / nt, k = sk % nt;
            all.insert(all.end(), {s, k, sk});
            if (k >= 1) {
                prop.insert(prop.end(), {s, sk - 1, sk});
            }
        }
        L.off_all[u + 1] = int(all.size() / 3);
        L.off_prop[u + 1] = int(prop.size() / 3);
        L.mmax = std::max(L.mmax, int(at[u].size()));
    }
    L.prop_base = int(all.size() / 3);
    L.data = all;
    L.data.insert(L.data.end(), prop.begin(), prop.end());
    if (L.data.empty()) L.data.push_back(0);
    return L;
}

// ---- kernels ----------------------------------------------------------------

// f_{s,0} = mu(t_s) (real)
template <typename T>
__global__ void __launch_bounds__(kAdjNT)
adj_init_kernel(int N, const int* __restrict__ trip, const T* __restrict__ S, T* __restrict__ fv) {
    const int i = blockIdx.x * kAdjNT + threadIdx.x;
    const int* t = trip + 3 * blockIdx.y;
    if (i >= N || t[1] != 0) return;
    T* f = fv + size_t(t[2]) * N * 6 + size_t(i) * 6;
    for (int a = 0; a < 3; ++a) {
        f[a] = S[(CIR_SITE_MU + a) * N + i];
        f[3 + a] = T(0);
    }
}

// Y (N x 6m, Re columns [0,3m), Im [3m,6m)) <-> blocks of fv / b.
// trip[3v + sel] names the block of column group v.
template <typename T, bool PACK>
__global__ void __launch_bounds__(kAdjNT)
adj_pack_kernel(int N, int m, const int* __restrict__ trip, int sel, int bstride, T* __restrict__ store,
                T* __restrict__ Y) {
    const int64_t e = int64_t(blockIdx.x) * kAdjNT + threadIdx.x;
    const int w6 = 6 * m;
    if (e >= int64_t(N) * w6) return;
    const int i = int(e / w6), c = int(e % w6);
    const int c3 = c % (3 * m);
    const int v = c3 / 3, a = c3 % 3, im = c >= 3 * m ? 3 : 0;
    T* s = store + size_t(trip[3 * v + sel]) * bstride + size_t(i) * 6 + im + a;
    if (PACK)
        Y[size_t(i) * w6 + c] = *s;
    else
        *s = Y[size_t(i) * w6 + c];
}

// C contribution of each pair at a frame: sum_a mu_a(t)^T f_{s,k,a}
template <typename T>
__global__ void __launch_bounds__(kAdjNT)
adj_dot_kernel(int N, const int* __restrict__ trip, const T* __restrict__ S, const T* __restrict__ fv,
               double* __restrict__ cpair) {
    const int* t = trip + 3 * blockIdx.x;
    const T* f = fv + size_t(t[2]) * N * 6;
    double re = 0.0, im = 0.0;
    for (int i = threadIdx.x; i < N; i += kAdjNT)
        for (int a = 0; a < 3; ++a) {
            const double mu = double(S[(CIR_SITE_MU + a) * N + i]);
            re += mu * double(f[size_t(i) * 6 + a]);
            im += mu * double(f[size_t(i) * 6 + 3 + a]);
        }
    __shared__ double red[2][kAdjNT / 64];
    re = wave_sum(re);
    im = wave_sum(im);
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = re;
        red[1][threadIdx.x >> 6] = im;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = 0.0, q = 0.0;
        for (int g = 0; g < kAdjNT / 64; ++g) {
            r += red[0][g];
            q += red[1][g];
        }
        cpair[2 * t[2]] = r;
        cpair[2 * t[2] + 1] = q;
    }
}

// b_s <- (k == nt-1 ? 0 : b_s) + w_k mu(t)
template <typename T>
__global__ void __launch_bounds__(kAdjNT)
adj_badd_kernel(int N, int nt, const int* __restrict__ trip, const double* __restrict__ wk, const T* __restrict__ S,
                T* __restrict__ b) {
    const int i = blockIdx.x * kAdjNT + threadIdx.x;
    const int* t = trip + 3 * blockIdx.y;
    if (i >= N) return;
    const int k = t[1];
    const bool first = k == nt - 1;
    const T wr = T(wk[2 * k]), wi = T(wk[2 * k + 1]);
    T* bb = b + size_t(t[0]) * N * 6 + size_t(i) * 6;
    for (int a = 0; a < 3; ++a) {
        const T mu = S[(CIR_SITE_MU + a) * N + i];
        bb[a] = (first ? T(0) : bb[a]) + wr * mu;
        bb[3 + a] = (first ? T(0) : bb[3 + a]) + wi * mu;
    }
}

// dL/dmu(t) = sum over pairs of Re(w_k f_{s,k}) (+ Re b_s at the sample's
// first point)
template <typename T>
__global__ void __launch_bounds__(kAdjNT)
adj_gmu_kernel(int N, int m, const int* __restrict__ trip, const double* __restrict__ wk, const T* __restrict__ fv,
               const T* __restrict__ b, double* __restrict__ gmu) {
    const int i = blockIdx.x * kAdjNT + threadIdx.x;
    if (i >= N) return;
    double g[3] = {0.0, 0.0, 0.0};
    for (int v = 0; v < m; ++v) {
        const int* t = trip + 3 * v;
        const int k = t[1];
        const double wr = wk[2 * k], wi = wk[2 * k + 1];
        const T* f = fv + size_t(t[2]) * N * 6 + size_t(i) * 6;
        const T* bb = b + size_t(t[0]) * N * 6 + size_t(i) * 6;
        for (int a = 0; a < 3; ++a) {
            g[a] += wr * double(f[a]) - wi * double(f[3 + a]);
            if (k == 0) g[a] += double(bb[a]);
        }
    }
    for (int a = 0; a < 3; ++a) gmu[size_t(a) * N + i] = g[a];
}

// R = Re(D o (P + P^T)) with P = Pr + i Pi (row-major, ld), tiles of 32x32
// staged through LDS for the transposed reads.
template <typename T>
__global__ void __launch_bounds__(256)
adj_rsym_kernel(int N, int ld, double tau, const double* __restrict__ lam, const T* __restrict__ Pr,
                const T* __restrict__ Pi, T* __restrict__ R) {
    __shared__ T tr[32][33], ti[32][33];
    const int k0 = blockIdx.y * 32, l0 = blockIdx.x * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    // transposed tile: P(l0 + r, k0 + c)
    for (int r = ty; r < 32; r += 8) {
        const int l = l0 + r, k = k0 + tx;
        const bool in = l < N && k < N;
        tr[r][tx] = in ? Pr[size_t(l) * ld + k] : T(0);
        ti[r][tx] = in ? Pi[size_t(l) * ld + k] : T(0);
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int k = k0 + r, l = l0 + tx;
        if (k >= N || l >= N) continue;
        const double sr = double(Pr[size_t(k) * ld + l]) + double(tr[tx][r]);
        const double si = double(Pi[size_t(k) * ld + l]) + double(ti[tx][r]);
        const double x = 0.5 * tau * (lam[k] - lam[l]);
        const double sc = fabs(x) < 1e-4 ? 1.0 - x * x / 6.0 : sin(x) / x;
        double sn, cs;
        sincos(0.5 * tau * (lam[k] + lam[l]), &sn, &cs);
        const double dr = -tau * sc * sn, di = tau * sc * cs;
        R[size_t(k) * ld + l] = T(dr * sr - di * si);
    }
}

// Contraction of the symmetric Gs = G + G^T of one frame with dH/dcoef, plus
// the dipole terms; per-workgroup partial sums of the 13 coefficient
// derivatives are added into part[blockIdx.x * 13 + c].
template <typename T>
__global__ void __launch_bounds__(256)
adj_coef_kernel(int N, int ld, int nchrom_mol, int hod, MapT<T> mc, const float* __restrict__ box, int boxk,
                int frame, const T* __restrict__ S, const T* __restrict__ Gs, const double* __restrict__ gmu,
                double* __restrict__ part) {
    __shared__ T cj[10][256];  // E, x, p, s, u(3), r(3) of a column tile
    __shared__ double red[kCoefRows][CIR_NCOEF];
    const int tid = threadIdx.x, lane = tid & 15, rl = tid >> 4;
    const int i = blockIdx.x * kCoefRows + rl;
    const bool act = i < N;
    const int ii = act ? i : 0;
    const T Ei = S[CIR_SITE_E * N + ii], wi = S[CIR_SITE_W * N + ii], xi = S[CIR_SITE_X * N + ii];
    const T pi = S[CIR_SITE_P * N + ii], mupi = S[CIR_SITE_MUP * N + ii];
    const T ui[3] = {S[CIR_SITE_U * N + ii], S[(CIR_SITE_U + 1) * N + ii], S[(CIR_SITE_U + 2) * N + ii]};
    const T ri[3] = {S[CIR_SITE_R * N + ii], S[(CIR_SITE_R + 1) * N + ii], S[(CIR_SITE_R + 2) * N + ii]};
    const PBox<T> bb = load_box<T>(box, boxk, frame);
    const int moli = ii / nchrom_mol;
    T gx = T(0), gp = T(0), gs = T(0), c0 = T(0), c1 = T(0), c2 = T(0);
    if (Gs != nullptr && !hod) {
        for (int j0 = 0; j0 < N; j0 += 256) {
            __syncthreads();
            const int j = j0 + tid;
            if (j < N) {
                cj[0][tid] = S[CIR_SITE_E * N + j];
                cj[1][tid] = S[CIR_SITE_X * N + j];
                cj[2][tid] = S[CIR_SITE_P * N + j];
                cj[3][tid] = S[CIR_SITE_X * N + j] * S[CIR_SITE_MUP * N + j];
                for (int c = 0; c < 3; ++c) {
                    cj[4 + c][tid] = S[(CIR_SITE_U + c) * N + j];
                    cj[7 + c][tid] = S[(CIR_SITE_R + c) * N + j];
                }
            }
            __syncthreads();
            if (!act) continue;
            const int cnt = min(256, N - j0);
            for (int jj = lane; jj < cnt; jj += 16) {
                const int jg = j0 + jj;
                if (jg == ii) continue;
                const T g = Gs[size_t(ii) * ld + jg];
                if (jg / nchrom_mol == moli) {
                    // calcIR.cu:1330-1341
                    const T xx = xi * cj[1][jj];
                    const T es = Ei + cj[0][jj];
                    gx += g * (mc.wi0 + mc.wi1 * es) * cj[1][jj];
                    gp += g * mc.wi2 * cj[2][jj];
                    c0 += g * xx;
                    c1 += g * es * xx;
                    c2 += g * (pi * cj[2][jj]);
                } else {
                    const T geo = tdc_geometry<T>(ui[0], ui[1], ui[2], ri[0], ri[1], ri[2], cj[4][jj], cj[5][jj],
                                                  cj[6][jj], cj[7][jj], cj[8][jj], cj[9][jj], bb);
                    gs += g * geo * cj[3][jj];
                }
            }
        }
    }
    // the 16 lanes of a row are consecutive lanes of one wave
    for (int o = 8; o > 0; o >>= 1) {
        gx += __shfl_xor(gx, o, 16);
        gp += __shfl_xor(gp, o, 16);
        gs += __shfl_xor(gs, o, 16);
        c0 += __shfl_xor(c0, o, 16);
        c1 += __shfl_xor(c1, o, 16);
        c2 += __shfl_xor(c2, o, 16);
    }
    if (lane == 0) {
        double c[CIR_NCOEF] = {};
        if (act) {
            const double E = double(Ei), W = double(wi);
            const double gii = Gs != nullptr ? 0.5 * double(Gs[size_t(ii) * ld + ii]) : 0.0;
            double gsd = double(gs) * kCmPerHartree;
            for (int a = 0; a < 3; ++a) gsd += gmu[size_t(a) * N + ii] * double(ui[a]);  // mu = u x mu'
            const double gxd = double(gx) + gsd * double(mupi);
            const double gm = gsd * double(xi);
            const double gpd = double(gp);
            const double gw = gii + double(mc.x1) * gxd + double(mc.p1) * gpd;
            c[0] = gw;
            c[1] = gw * E;
            c[2] = gw * E * E;
            c[3] = gxd;
            c[4] = gxd * W;
            c[5] = gpd;
            c[6] = gpd * W;
            c[7] = gm;
            c[8] = gm * E;
            c[9] = gm * E * E;
            c[10] = 0.5 * double(c0);
            c[11] = 0.5 * double(c1);
            c[12] = 0.5 * double(c2);
        }
        for (int q = 0; q < CIR_NCOEF; ++q) red[rl][q] = c[q];
    }
    __syncthreads();
    if (tid < CIR_NCOEF) {
        double acc = 0.0;
        for (int r = 0; r < kCoefRows; ++r) acc += red[r][tid];
        part[size_t(blockIdx.x) * CIR_NCOEF + tid] += acc;
    }
}

// ---- host helpers -------------------------------------------------------------

template <typename T>
int pack(bool to_y, int N, int m, const int* trip, int sel, void* store, void* Y, hipStream_t s) {
    if (m <= 0) return CIR_OK;
    const int64_t tot = int64_t(N) * 6 * m;
    const unsigned g = unsigned((tot + kAdjNT - 1) / kAdjNT);
    if (to_y)
        hipLaunchKernelGGL((adj_pack_kernel<T, true>), dim3(g), dim3(kAdjNT), 0, s, N, m, trip, sel, N * 6,
                           static_cast<T*>(store), static_cast<T*>(Y));
    else
        hipLaunchKernelGGL((adj_pack_kernel<T, false>), dim3(g), dim3(kAdjNT), 0, s, N, m, trip, sel, N * 6,
                           static_cast<T*>(store), static_cast<T*>(Y));
    CIR_CHECK_LAUNCH();
    return CIR_OK;
}

int pack_any(int rp, bool to_y, int N, int m, const int* trip, int sel, void* store, void* Y, hipStream_t s) {
    return rp == CIR_PREC_FP64 ? pack<double>(to_y, N, m, trip, sel, store, Y, s)
                               : pack<float>(to_y, N, m, trip, sel, store, Y, s);
}

GemmArgs rowmajor(int M, int Nn, int K, const void* A, int64_t a_m, int64_t a_k, const void* B, int64_t b_k,
                  int64_t b_n, void* C, int64_t c_m, double alpha, double beta) {
    GemmArgs g;
    g.M = M;
    g.N = Nn;
    g.K = K;
    g.A = A;
    g.a_m = a_m;
    g.a_k = a_k;
    g.B = B;
    g.b_k = b_k;
    g.b_n = b_n;
    g.C = C;
    g.c_m = c_m;
    g.c_n = 1;
    g.alpha = alpha;
    g.beta = beta;
    return g;
}

template <typename T>
int forward_impl(cir_ctx* c, int nframes, const float* xyz_d, const float* box_d, const float* times,
                 double* ftcf) {
    const cir_params& p = c->p;
    std::unique_ptr<AdjState, AdjDeleter> A(new AdjState());
    CIR_TRY(build_schedule(p, nframes, times, A->S));
    const Schedule& S = A->S;
    const int nu = int(S.uniq.size());
    const int N = c->N, ld = c->ld, ns = p.nsamples, nt = p.ntcfpoints;
    const int rp = c->real_prec;
    const size_t esz = sizeof(T);
    hipStream_t st = c->stream;
    PairLists L = build_pairs(p, S);
    A->nu = nu;
    A->mmax = L.mmax;
    A->off_all = L.off_all;
    A->off_prop = L.off_prop;
    for (int& o : A->off_prop) o += L.prop_base;
    // HBM held between the calls: sites of every distinct frame, the dipole
    // vectors of every (sample, point) and either every frame's eigensystem
    // (store) or one chunk's, recomputed chunk by chunk in the reverse sweep
    const int B = auto_chunk(c, nu);
    const size_t base = size_t(nu) * size_t(CIR_NSITE) * N * esz + size_t(ns) * nt * N * 6 * esz +
                        size_t(B) * N * ld * esz + eigh_workspace(rp, B, N, ld);
    const size_t eig_all = size_t(nu) * (size_t(N) * ld * esz + size_t(N) * 8);
    size_t fr = 0, tot = 0;
    CIR_HIP(hipMemGetInfo(&fr, &tot));
    const bool fits = double(base + eig_all) <= 0.8 * double(fr);
    if (p.adj_store == 1 && !fits)
        return fail(CIR_ERR_NOMEM, "forward_ls: the eigensystems of all frames do not fit in device memory");
    A->store = p.adj_store == 1 || (p.adj_store == 0 && fits);
    A->B = B;
    const int nvs = A->store ? nu : B;  // frames held in V, w
    CIR_TRY(A->V.alloc(size_t(nvs) * N * ld * esz));
    CIR_TRY(A->w.alloc(size_t(nvs) * N * sizeof(double)));
    CIR_TRY(A->H.alloc(size_t(B) * N * ld * esz));
    CIR_TRY(A->work.alloc(eigh_workspace(rp, B, N, ld)));
    CIR_TRY(A->sites.alloc(size_t(nu) * CIR_NSITE * N * esz));
    CIR_TRY(A->box.alloc(size_t(nu) * box_floats(c) * sizeof(float)));
    CIR_TRY(A->fv.alloc(size_t(ns) * nt * N * 6 * esz));
    CIR_TRY(A->pairs.alloc(L.data.size() * sizeof(int)));
    CIR_HIP(hipMemcpyAsync(A->pairs.p, L.data.data(), L.data.size() * sizeof(int), hipMemcpyHostToDevice, st));
    const int* trip = A->pairs.as<int>();

    DevBuf xyzc, Yin, Yout, G, cpair;
    CIR_TRY(xyzc.alloc(size_t(B) * c->natoms * 3 * sizeof(float)));
    CIR_TRY(Yin.alloc(size_t(N) * 6 * L.mmax * esz));
    CIR_TRY(Yout.alloc(size_t(N) * 6 * L.mmax * esz));
    CIR_TRY(G.alloc(prop_work_max(N, L.mmax) * esz));
    CIR_TRY(cpair.alloc(size_t(ns) * nt * 2 * sizeof(double)));
    char* Vb = A->V.as<char>();
    char* Sb = A->sites.as<char>();
    float* boxes = A->box.as<float>();
    for (int c0 = 0; c0 < nu; c0 += B) {
        const int nb = std::min(B, nu - c0);
        const int bs = box_floats(c);
        CIR_TRY(gather_frames(c, xyz_d, box_d, S.uniq.data() + c0, nb, xyzc.as<float>(), boxes + bs * size_t(c0)));
        void* sc = Sb + size_t(c0) * CIR_NSITE * N * esz;
        CIR_TRY(eproj_sites(rp, nb, c->natoms, p.natom_mol, p.nchrom_mol, c->imodel, site_rules(p), c->mc, p.cutoff,
                            xyzc.as<float>(), boxes + bs * size_t(c0), c->boxk, sc, st));
        const int v0 = A->store ? c0 : 0;  // slot of frame c0 in V, w
        CIR_TRY(hamiltonian(rp, nb, N, ld, p.nchrom_mol, c->ispecies, c->mc, p.avef, boxes + bs * size_t(c0),
                            c->boxk, sc, A->H.p, st));
        CIR_TRY(eigh_batched(c->prop_prec, nb, N, ld, A->H.p, Vb + size_t(v0) * N * ld * esz,
                             A->w.as<double>() + size_t(v0) * N, A->work.p, st, nullptr));
        for (int u = c0; u < c0 + nb; ++u) {
            const int vu = u - c0 + v0;
            const T* Su = reinterpret_cast<const T*>(Sb + size_t(u) * CIR_NSITE * N * esz);
            const int a0 = L.off_all[u], na = L.off_all[u + 1] - a0;
            const int p0 = A->off_prop[u], np = A->off_prop[u + 1] - p0;
            if (na == 0) continue;
            hipLaunchKernelGGL(adj_init_kernel<T>, dim3(unsigned(ceil_div(N, kAdjNT)), unsigned(na)), dim3(kAdjNT), 0,
                               st, N, trip + 3 * a0, Su, A->fv.as<T>());
            CIR_CHECK_LAUNCH();
            if (np > 0) {
                CIR_TRY(pack_any(rp, true, N, np, trip + 3 * p0, 1, A->fv.p, Yin.p, st));
                CIR_TRY(propagate(rp, c->prop_prec, N, ld, 3 * np, 6 * np, double(p.dt),
                                  Vb + size_t(vu) * N * ld * esz, A->w.as<double>() + size_t(vu) * N, Yin.p, Yout.p,
                                  G.p, st));
                CIR_TRY(pack_any(rp, false, N, np, trip + 3 * p0, 2, A->fv.p, Yout.p, st));
            }
            hipLaunchKernelGGL(adj_dot_kernel<T>, dim3(unsigned(na)), dim3(kAdjNT), 0, st, N, trip + 3 * a0, Su,
                               A->fv.as<T>(), cpair.as<double>());
            CIR_CHECK_LAUNCH();
        }
    }
    std::vector<double> cp(size_t(ns) * nt * 2);
    CIR_HIP(hipMemcpyAsync(cp.data(), cpair.p, cp.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    CIR_HIP(hipStreamSynchronize(st));
    // calcIR.cu:646-648 (sum over samples), 777-781 (decay), 788-827 (C2R)
    A->tcf.assign(2 * size_t(nt), 0.0);
    std::vector<double> X(2 * size_t(nt));
    for (int k = 0; k < nt; ++k) {
        for (int s = 0; s < ns; ++s) {
            A->tcf[2 * k] += cp[(size_t(s) * nt + k) * 2];
            A->tcf[2 * k + 1] += cp[(size_t(s) * nt + k) * 2 + 1];
        }
        const double dcy = std::exp(-1.0 * k * p.dt / (2.0 * p.t1)) / (1. * ns);
        X[2 * k] = A->tcf[2 * k] * dcy;
        X[2 * k + 1] = A->tcf[2 * k + 1] * dcy;
    }
    DevBuf dX, dy;
    CIR_TRY(dX.alloc(X.size() * sizeof(double)));
    CIR_TRY(dy.alloc(size_t(c->nfft) * sizeof(double)));
    CIR_HIP(hipMemcpyAsync(dX.p, X.data(), X.size() * sizeof(double), hipMemcpyHostToDevice, st));
    CIR_TRY(spectrum_c2r(1, c->nfft, nt, dX.as<double>(), dy.as<double>(), st));
    CIR_HIP(hipMemcpyAsync(ftcf, dy.p, size_t(c->nfft) * sizeof(double), hipMemcpyDeviceToHost, st));
    CIR_HIP(hipStreamSynchronize(st));
    A->have = true;
    c->adj = std::move(A);
    return CIR_OK;
}

template <typename T>
int backward_impl(cir_ctx* c, const double* dL_dy, double* grad) {
    AdjState* A = c->adj.get();
    const cir_params& p = c->p;
    const int N = c->N, ld = c->ld, ns = p.nsamples, nt = p.ntcfpoints, nfft = c->nfft, nu = A->nu;
    const int rp = c->real_prec;
    const size_t esz = sizeof(T);
    hipStream_t st = c->stream;
    // weights w_k = dL/dRe C_k - i dL/dIm C_k: the transpose of the C2R
    // (c2r_kernel) and of the decay / 1/nsamples scaling
    std::vector<double> wk(2 * size_t(nt), 0.0);
    const int half = nfft / 2, jmax = std::min(nt - 1, half);
    for (int j = 0; j < nt; ++j) {
        double gr = 0.0, gi = 0.0;
        if (j == 0) {
            for (int m = 0; m < nfft; ++m) gr += dL_dy[m];
        } else if (j <= jmax) {
            for (int m = 0; m < nfft; ++m) {
                const double th = kTwoPiExact * double((int64_t(j) * m) % nfft) / double(nfft);
                if (nfft % 2 == 0 && j == half) {
                    gr += dL_dy[m] * std::cos(th);
                } else {
                    gr += 2.0 * dL_dy[m] * std::cos(th);
                    gi -= 2.0 * dL_dy[m] * std::sin(th);
                }
            }
        }
        const double dcy = std::exp(-1.0 * j * p.dt / (2.0 * p.t1)) / (1. * ns);
        wk[2 * j] = gr * dcy;
        wk[2 * j + 1] = -gi * dcy;
    }
    const int m = A->mmax;
    const int nrb = ceil_div(N, kCoefRows);
    DevBuf dwk, b, Bp, Fp, Bo, beta, phi, X1, X2, X3, G, gmu, part;
    CIR_TRY(dwk.alloc(wk.size() * sizeof(double)));
    CIR_TRY(b.alloc(size_t(ns) * N * 6 * esz));
    CIR_TRY(Bp.alloc(size_t(N) * 6 * m * esz));
    CIR_TRY(Fp.alloc(size_t(N) * 6 * m * esz));
    CIR_TRY(Bo.alloc(size_t(N) * 6 * m * esz));
    CIR_TRY(beta.alloc(size_t(N) * 6 * m * esz));
    CIR_TRY(phi.alloc(size_t(N) * 6 * m * esz));
    CIR_TRY(X1.alloc(size_t(N) * ld * esz));
    CIR_TRY(X2.alloc(size_t(N) * ld * esz));
    CIR_TRY(X3.alloc(size_t(N) * ld * esz));
    CIR_TRY(G.alloc(prop_work_max(N, m) * esz));
    CIR_TRY(gmu.alloc(3 * size_t(N) * sizeof(double)));
    CIR_TRY(part.alloc(size_t(nrb) * CIR_NCOEF * sizeof(double)));
    CIR_HIP(hipMemcpyAsync(dwk.p, wk.data(), wk.size() * sizeof(double), hipMemcpyHostToDevice, st));
    CIR_HIP(hipMemsetAsync(part.p, 0, part.bytes, st));
    CIR_HIP(hipMemsetAsync(b.p, 0, b.bytes, st));
    const int* trip = A->pairs.as<int>();
    const char* Vb = A->V.as<char>();
    const char* Sb = A->sites.as<char>();
    const double tau = double(p.dt) / kHbar;
    const int hod = (c->ispecies == 1 || c->ispecies == 2) ? 1 : 0;
    const MapT<T> mc = map_t<T>(c->mc);
    const int gp = rp;  // GEMM precision of the derivative products
    int cbeg = nu;  // first frame of the chunk whose eigensystems V, w hold (recompute mode)
    for (int u = nu - 1; u >= 0; --u) {
        const int a0 = A->off_all[u], na = A->off_all[u + 1] - a0;
        const int p0 = A->off_prop[u], np = A->off_prop[u + 1] - p0;
        if (na == 0) continue;
        if (!A->store && u < cbeg) {
            // the chunk [cbeg, cbeg + nb) holding u, as the forward chunked it
            cbeg = (u / A->B) * A->B;
            const int nb = std::min(A->B, nu - cbeg);
            const void* sc = Sb + size_t(cbeg) * CIR_NSITE * N * esz;
            CIR_TRY(hamiltonian(rp, nb, N, ld, p.nchrom_mol, c->ispecies, c->mc, p.avef,
                                A->box.as<float>() + box_floats(c) * size_t(cbeg), c->boxk, sc, A->H.p, st));
            CIR_TRY(eigh_batched(c->prop_prec, nb, N, ld, A->H.p, A->V.p, A->w.as<double>(), A->work.p, st, nullptr));
        }
        const int vu = A->store ? u : u - cbeg;
        const T* Su = reinterpret_cast<const T*>(Sb + size_t(u) * CIR_NSITE * N * esz);
        const void* Vu = Vb + size_t(vu) * N * ld * esz;
        const double* wu = A->w.as<double>() + size_t(vu) * N;
        hipLaunchKernelGGL(adj_badd_kernel<T>, dim3(unsigned(ceil_div(N, kAdjNT)), unsigned(na)), dim3(kAdjNT), 0, st,
                           N, nt, trip + 3 * a0, dwk.as<double>(), Su, b.as<T>());
        CIR_CHECK_LAUNCH();
        hipLaunchKernelGGL(adj_gmu_kernel<T>, dim3(unsigned(ceil_div(N, kAdjNT))), dim3(kAdjNT), 0, st, N, na,
                           trip + 3 * a0, dwk.as<double>(), A->fv.as<T>(), b.as<T>(), gmu.as<double>());
        CIR_CHECK_LAUNCH();
        const T* Gs = nullptr;
        if (np > 0) {
            const int w6 = 6 * np, w3 = 3 * np;
            // gather b_k (by sample) and f_{k-1}
            {
                const int64_t tot = int64_t(N) * w6;
                const unsigned g = unsigned((tot + kAdjNT - 1) / kAdjNT);
                hipLaunchKernelGGL((adj_pack_kernel<T, true>), dim3(g), dim3(kAdjNT), 0, st, N, np, trip + 3 * p0, 0,
                                   N * 6, b.as<T>(), Bp.as<T>());
                CIR_CHECK_LAUNCH();
            }
            CIR_TRY(pack_any(rp, true, N, np, trip + 3 * p0, 1, A->fv.p, Fp.p, st));
            // beta = V^T b, phi = V^T f (V(i,k) at V[i*ld+k])
            CIR_TRY(gemm(gp, rowmajor(N, w6, N, Vu, 1, ld, Bp.p, w6, 1, beta.p, w6, 1.0, 0.0), st));
            CIR_TRY(gemm(gp, rowmajor(N, w6, N, Vu, 1, ld, Fp.p, w6, 1, phi.p, w6, 1.0, 0.0), st));
            const char* be = beta.as<char>();
            const char* ph = phi.as<char>();
            const size_t im = size_t(w3) * esz;
            // Pr = br fr^T - bi fi^T, Pi = br fi^T + bi fr^T (X1, X2)
            CIR_TRY(gemm(gp, rowmajor(N, N, w3, be, w6, 1, ph, 1, w6, X1.p, ld, 1.0, 0.0), st));
            CIR_TRY(gemm(gp, rowmajor(N, N, w3, be + im, w6, 1, ph + im, 1, w6, X1.p, ld, -1.0, 1.0), st));
            CIR_TRY(gemm(gp, rowmajor(N, N, w3, be, w6, 1, ph + im, 1, w6, X2.p, ld, 1.0, 0.0), st));
            CIR_TRY(gemm(gp, rowmajor(N, N, w3, be + im, w6, 1, ph, 1, w6, X2.p, ld, 1.0, 1.0), st));
            hipLaunchKernelGGL(adj_rsym_kernel<T>, dim3(unsigned(ceil_div(N, 32)), unsigned(ceil_div(N, 32))),
                               dim3(256), 0, st, N, ld, tau, wu, X1.as<T>(), X2.as<T>(), X3.as<T>());
            CIR_CHECK_LAUNCH();
            // Gs = V R V^T: X1 = V R, X2 = X1 V^T
            CIR_TRY(gemm(gp, rowmajor(N, N, N, Vu, ld, 1, X3.p, ld, 1, X1.p, ld, 1.0, 0.0), st));
            CIR_TRY(gemm(gp, rowmajor(N, N, N, X1.p, ld, 1, Vu, 1, ld, X2.p, ld, 1.0, 0.0), st));
            Gs = X2.as<T>();
            // b <- U(t) b for the next (earlier) frame
            CIR_TRY(propagate(rp, c->prop_prec, N, ld, w3, w6, double(p.dt), Vu, wu, Bp.p, Bo.p, G.p, st));
            {
                const int64_t tot = int64_t(N) * w6;
                const unsigned g = unsigned((tot + kAdjNT - 1) / kAdjNT);
                hipLaunchKernelGGL((adj_pack_kernel<T, false>), dim3(g), dim3(kAdjNT), 0, st, N, np, trip + 3 * p0,
                                   0, N * 6, b.as<T>(), Bo.as<T>());
                CIR_CHECK_LAUNCH();
            }
        }
        hipLaunchKernelGGL(adj_coef_kernel<T>, dim3(unsigned(nrb)), dim3(256), 0, st, N, ld, p.nchrom_mol, hod, mc,
                           A->box.as<float>(), c->boxk, u, Su, Gs, gmu.as<double>(), part.as<double>());
        CIR_CHECK_LAUNCH();
    }
    std::vector<double> hp(size_t(nrb) * CIR_NCOEF);
    CIR_HIP(hipMemcpyAsync(hp.data(), part.p, hp.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    CIR_HIP(hipStreamSynchronize(st));
    for (int q = 0; q < CIR_NCOEF; ++q) {
        double acc = 0.0;
        for (int r = 0; r < nrb; ++r) acc += hp[size_t(r) * CIR_NCOEF + q];
        grad[q] = acc;
    }
    return CIR_OK;
}

void coef_to_array(const MapCoef& m, double* c) {
    const double v[CIR_NCOEF] = {m.w[0], m.w[1], m.w[2], m.x[0], m.x[1], m.p[0], m.p[1],
                                 m.mup[0], m.mup[1], m.mup[2], m.wi[0], m.wi[1], m.wi[2]};
    std::memcpy(c, v, sizeof(v));
}

MapCoef array_to_coef(const double* c) {
    return MapCoef{{c[0], c[1], c[2]}, {c[3], c[4]}, {c[5], c[6]}, {c[7], c[8], c[9]}, {c[10], c[11], c[12]}};
}

}  // namespace
}  // namespace cir

using namespace cir;

extern "C" int cir_map_coefs(int imap, int ispecies, double* coef) {
    if (!coef) return fail(CIR_ERR_ARG, "cir_map_coefs: null argument");
    MapCoef mc;
    if (!map_coef(imap, ispecies, &mc)) return fail(CIR_ERR_MAP, "map/species not in the map table");
    coef_to_array(mc, coef);
    return CIR_OK;
}

extern "C" int cir_get_coefs(const cir_ctx* c, double* coef) {
    if (!c || !coef) return fail(CIR_ERR_ARG, "cir_get_coefs: null argument");
    coef_to_array(c->mc, coef);
    return CIR_OK;
}

extern "C" int cir_set_coefs(cir_ctx* c, const double* coef) {
    if (!c) return fail(CIR_ERR_ARG, "cir_set_coefs: null context");
    if (coef == nullptr) {
        if (!map_coef(c->p.imap, c->ispecies, &c->mc)) return fail(CIR_ERR_MAP, "map/species not in the map table");
    } else {
        for (int q = 0; q < CIR_NCOEF; ++q)
            if (!std::isfinite(coef[q])) return fail(CIR_ERR_ARG, "cir_set_coefs: non-finite coefficient");
        c->mc = array_to_coef(coef);
    }
    c->adj.reset();
    c->have = false;
    return CIR_OK;
}

extern "C" int cir_forward_ls(cir_ctx* c, int nframes, const float* xyz, const float* box, const float* times,
                              double* ftcf_ir) try {
    if (!c || !xyz || !box || !times || !ftcf_ir || nframes <= 0)
        return fail(CIR_ERR_ARG, "cir_forward_ls: bad arguments");
    if (c->p.adj_store < 0 || c->p.adj_store > 2) return fail(CIR_ERR_ARG, "adjoint_store must be 0, 1 or 2");
    if (c->p.reunit_every > 0)
        return fail(CIR_ERR_ARG, "the lineshape adjoint propagates without re-unitarisation: set reunitarise_every 0");
    c->adj.reset();
    const size_t fb = size_t(nframes) * c->natoms * 3 * sizeof(float);
    for (size_t q = 0; q < size_t(nframes) * c->natoms * 3; ++q)
        if (!std::isfinite(xyz[q])) return fail(CIR_ERR_NONFINITE, "non-finite coordinate");
    CIR_TRY(check_boxes(c, nframes, box));
    const size_t nbx = size_t(nframes) * box_floats(c);
    DevBuf dx, db;
    CIR_TRY(dx.alloc(fb));
    CIR_TRY(db.alloc(nbx * sizeof(float)));
    CIR_HIP(hipMemcpyAsync(dx.p, xyz, fb, hipMemcpyHostToDevice, c->stream));
    CIR_HIP(hipMemcpyAsync(db.p, box, nbx * sizeof(float), hipMemcpyHostToDevice, c->stream));
    if (c->real_prec == CIR_PREC_FP64)
        return forward_impl<double>(c, nframes, dx.as<float>(), db.as<float>(), times, ftcf_ir);
    return forward_impl<float>(c, nframes, dx.as<float>(), db.as<float>(), times, ftcf_ir);
} catch (...) {
    return cir::host_exception("cir_forward_ls");
}

extern "C" int cir_backward_ls(cir_ctx* c, const double* dL_dftcf_ir, double* dL_dcoef) try {
    if (!c || !dL_dftcf_ir || !dL_dcoef) return fail(CIR_ERR_ARG, "cir_backward_ls: bad arguments");
    if (!c->adj || !c->adj->have) return fail(CIR_ERR_ARG, "cir_backward_ls: call cir_forward_ls first");
    if (c->real_prec == CIR_PREC_FP64) return backward_impl<double>(c, dL_dftcf_ir, dL_dcoef);
    return backward_impl<float>(c, dL_dftcf_ir, dL_dcoef);
} catch (...) {
    return cir::host_exception("cir_backward_ls");
}
