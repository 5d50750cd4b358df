// oracle.cpp -- CPU restatement of calcIR's spectroscopy pipeline.
//
// TEST INFRASTRUCTURE ONLY (see oracle.h).  The product never links this
// file.  It is compiled with -ffp-contract=off so that the FP32 cutoff test
// below is evaluated with the exact IEEE operation sequence of the reference
// (calcIR.cu:1106-1113) and of the HIP field kernel.
//
// Every routine is a template over the working real type T (float = the
// reference's user_real_t, calcIR.h:22; double = the error reference).

#include "oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <limits>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// calcIR.h:36-37
constexpr double kHbar = 5.308837367;      // cm^-1 ps
constexpr double kPi = 3.14159265359;
// calcIR.cu:1036, 1197, 1198
constexpr double kBohrPerNm = 18.8973;
constexpr double kCmPerHartree = 2.1947463e5;
// calcIR.cu:1136-1138 (TIP4P O-M distance), 1289 (TDM offset), 1299 (Raman)
constexpr double kOMTip4p = 0.0150;
constexpr double kTdmOffset = 0.067;
constexpr double kRamanAniso = 4.6;
// calcIR.cu:1152, 1157
constexpr double kQH = 0.52;
constexpr double kQM = 1.04;
// The C2R transform (cuFFT's contract) uses the exact value of 2*pi.
constexpr double kTwoPiExact = 6.283185307179586476925286766559;

// Spectroscopic map coefficients, calcIR.cu:1203-1242.
struct MapCoef {
    double w[3], x[2], p[2], mup[3], wi[3];
};

bool map_table(int imap, int ispecies, MapCoef* c) {
    if (ispecies < 0 || ispecies > 3) return false;
    const bool light = (ispecies == 0 || ispecies == 1);  // H2O, HOD/D2O
    if (imap == 0) {  // Gruenbaum 2013
        if (light) {
            *c = {{3670.2, -3541.7, -152677.0}, {0.19285, -1.7261e-5}, {1.6466, 5.7692e-4},
                  {0.1646, 11.39, 63.41}, {-1361.0, 27165.0, -1.887}};
        } else {
            *c = {{2767.8, -2630.3, -102601.0}, {0.16593, -2.0632e-5}, {2.0475, 8.9108e-4},
                  {0.1646, 11.39, 63.41}, {-1361.0, 27165.0, -1.887}};
        }
        return true;
    }
    if (imap == 1) {  // Li and Skinner 2010
        if (light) {
            *c = {{3732.9, -3519.8, -153520.0}, {0.19318, -1.7248e-5}, {1.6102, 5.8697e-4},
                  {0.1622, 10.381, 137.6}, {-1360.8, 27171.0, -1.887}};
        } else {
            *c = {{2748.2, -2572.2, -102980.0}, {0.16598, -2.0752e-5}, {1.9813, 9.1419e-4},
                  {0.1622, 10.381, 137.6}, {-1360.8, 27171.0, -1.887}};
        }
        return true;
    }
    return false;
}

// The table row, or the caller's coefficients when the config carries them.
bool cfg_coef(const orc_config* cfg, MapCoef* c) {
    if (!map_table(cfg->imap, cfg->ispecies, c)) return false;
    if (cfg->use_coef) {
        const double* k = cfg->coef;
        *c = {{k[0], k[1], k[2]}, {k[3], k[4]}, {k[5], k[6]}, {k[7], k[8], k[9]}, {k[10], k[11], k[12]}};
    }
    return true;
}

// calcIR.cu:1457-1460: dx - L*round(dx/L), round half away from zero.
template <typename T>
inline T minimage(T dx, T boxl) {
    return dx - boxl * std::round(dx / boxl);
}

// Periodic box of a frame.  Rectangular (the reference, calcIR.cu:404-405):
// the diagonal, 3 floats per frame.  With cfg->triclinic: 9 floats per frame,
// the rows a = (a0,0,0), b = (b0,b1,0), c = (c0,c1,c2) of a GROMACS box.
template <typename T>
struct Box {
    T a0, b0, b1, c0, c1, c2;
    bool tri;
};

template <typename T>
Box<T> make_box(const orc_config* cfg, const float* q) {
    Box<T> B;
    if (!cfg->triclinic) {
        B.a0 = T(q[0]);
        B.b1 = T(q[1]);
        B.c2 = T(q[2]);
        B.b0 = B.c0 = B.c1 = T(0);
        B.tri = false;
    } else {
        B.a0 = T(q[0]);
        B.b0 = T(q[3]);
        B.b1 = T(q[4]);
        B.c0 = T(q[6]);
        B.c1 = T(q[7]);
        B.c2 = T(q[8]);
        B.tri = q[3] != 0.0f || q[6] != 0.0f || q[7] != 0.0f;
    }
    return B;
}

inline int box_floats(const orc_config* cfg) { return cfg->triclinic ? 9 : 3; }

// Minimum image of d.  Rectangular: per axis as calcIR.cu:1457-1460.
// Triclinic: remove c, then b, then a, each by the rounded coordinate along
// its own axis (the GROMACS form for reduced boxes).
template <typename T>
void mi3(T* d, const Box<T>& B) {
    if (!B.tri) {
        d[0] = minimage<T>(d[0], B.a0);
        d[1] = minimage<T>(d[1], B.b1);
        d[2] = minimage<T>(d[2], B.c2);
        return;
    }
    T s = std::round(d[2] / B.c2);
    d[0] = d[0] - s * B.c0;
    d[1] = d[1] - s * B.c1;
    d[2] = d[2] - s * B.c2;
    s = std::round(d[1] / B.b1);
    d[0] = d[0] - s * B.b0;
    d[1] = d[1] - s * B.b1;
    s = std::round(d[0] / B.a0);
    d[0] = d[0] - s * B.a0;
}

// The cutoff decision of calcIR.cu:1106-1113 is a discrete selection.  It is
// always evaluated in FP32 from the FP32 input coordinates, with the
// reference's operation order, so that every precision form selects the same
// molecules and precision comparisons are not polluted by selection flips.
inline bool within_cutoff_f32(const float* mo, const float* nh, const Box<float>& B, float cutoff) {
    float d[3] = {mo[0] - nh[0], mo[1] - nh[1], mo[2] - nh[2]};
    mi3<float>(d, B);
    float r = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    return !(r > cutoff);
}

// The H atom of a chromophore: calcIR.cu:1061-1070 (and 1266-1275,
// 1360-1370) takes H1 or H2 by chrom % 2 whatever nchrom_mol is, so with
// nchrom_mol 1 molecule n uses H1 for even n and H2 for odd n.  The
// oh_per_molecule variant takes H(1 + chrom % nchrom_mol) instead; the two
// agree for nchrom_mol 2.
inline int h_atom(const orc_config* cfg, int chrom) {
    return 1 + (cfg->oh_per_molecule ? chrom % cfg->nchrom_mol : chrom % 2);
}

// Layouts the reference's indexing supports: atoms 1 and 2 of every
// molecule are its hydrogens (calcIR.cu:1061-1070), so natom_mol >= 3; the
// oh_per_molecule variant names at most two OH stretches per molecule.
bool layout_ok(const orc_config* cfg, int natoms) {
    if (cfg->natom_mol < 3 || cfg->nchrom_mol <= 0 || natoms <= 0 || natoms % cfg->natom_mol) return false;
    return !(cfg->oh_per_molecule && cfg->nchrom_mol > 2);
}

int threads_of(const orc_config* cfg) {
#ifdef _OPENMP
    return cfg && cfg->nthreads > 0 ? cfg->nthreads : omp_get_max_threads();
#else
    (void)cfg;
    return 1;
#endif
}

// ---------------------------------------------------------------------------
// Electric field projection, calcIR.cu:1020-1168
// ---------------------------------------------------------------------------
template <typename T>
void eproj_t(const orc_config* cfg, const float* x, int natoms, const float* boxf, double* out) {
    const int nam = cfg->natom_mol, ncm = cfg->nchrom_mol;
    const int nmol = natoms / nam, nchrom = nmol * ncm;
    const Box<T> box = make_box<T>(cfg, boxf);
    const Box<float> boxs = make_box<float>(cfg, boxf);
    const T bohr = T(kBohrPerNm);
#pragma omp parallel for schedule(static) num_threads(threads_of(cfg))
    for (int chrom = 0; chrom < nchrom; ++chrom) {
        const int n = chrom / ncm;
        const float* nhf = &x[3 * (n * nam + h_atom(cfg, chrom))];
        const float* nof = &x[3 * (n * nam)];
        T nh[3] = {T(nhf[0]), T(nhf[1]), T(nhf[2])};
        T oh[3];
        for (int c = 0; c < 3; ++c) oh[c] = nh[c] - T(nof[c]);
        mi3<T>(oh, box);
        T r = std::sqrt(oh[0] * oh[0] + oh[1] * oh[1] + oh[2] * oh[2]);
        for (int c = 0; c < 3; ++c) oh[c] /= r;

        T ef[3] = {0, 0, 0};
        for (int m = 0; m < nmol; ++m) {
            if (m == n) continue;  // calcIR.cu:1099
            const float* mof = &x[3 * (m * nam)];
            if (!within_cutoff_f32(mof, nhf, boxs, cfg->cutoff)) continue;
            T mo[3] = {T(mof[0]), T(mof[1]), T(mof[2])};
            // calcIR.cu:1117-1160: atoms 1..natom_mol-1 of molecule m; atoms 1
            // and 2 carry +0.52, atom 3 carries -1.04, later atoms nothing.
            // The msite_bisector variant gives three-site water a -1.04 site
            // on the HOH bisector 0.015 nm from O (the TIP4P geometry).
            const bool built_m = cfg->msite_bisector && nam == 3;
            const int last = built_m ? 3 : std::min(nam - 1, 3);
            for (int i = 1; i <= last; ++i) {
                T mx[3] = {T(0), T(0), T(0)};
                if (i < nam) {
                    const float* maf = &x[3 * (m * nam + i)];
                    for (int c = 0; c < 3; ++c) mx[c] = T(maf[c]);
                }
                if (i == 3 && !built_m && cfg->imodel != 0) {
                    // calcIR.cu:1124-1140: M moved to 0.015 nm along O->M
                    T om[3];
                    for (int c = 0; c < 3; ++c) om[c] = mx[c] - mo[c];
                    mi3<T>(om, box);
                    T rom = std::sqrt(om[0] * om[0] + om[1] * om[1] + om[2] * om[2]);
                    for (int c = 0; c < 3; ++c) mx[c] = mo[c] + T(kOMTip4p) * om[c] / rom;
                } else if (i == 3 && built_m) {
                    T h1[3], h2[3], bi[3];
                    for (int c = 0; c < 3; ++c) {
                        h1[c] = T(x[3 * (m * nam + 1) + c]) - mo[c];
                        h2[c] = T(x[3 * (m * nam + 2) + c]) - mo[c];
                    }
                    mi3<T>(h1, box);
                    mi3<T>(h2, box);
                    const T r1 = std::sqrt(h1[0] * h1[0] + h1[1] * h1[1] + h1[2] * h1[2]);
                    const T r2 = std::sqrt(h2[0] * h2[0] + h2[1] * h2[1] + h2[2] * h2[2]);
                    for (int c = 0; c < 3; ++c) bi[c] = h1[c] / r1 + h2[c] / r2;
                    const T rb = std::sqrt(bi[0] * bi[0] + bi[1] * bi[1] + bi[2] * bi[2]);
                    for (int c = 0; c < 3; ++c) mx[c] = mo[c] + T(kOMTip4p) * bi[c] / rb;
                }
                T dr[3];
                for (int c = 0; c < 3; ++c) dr[c] = nh[c] - mx[c];
                mi3<T>(dr, box);
                for (int c = 0; c < 3; ++c) dr[c] = dr[c] * bohr;
                T rr = std::sqrt(dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2]);
                T r3 = rr * rr * rr;
                if (i < 3) {
                    for (int c = 0; c < 3; ++c) ef[c] += T(kQH) * dr[c] / r3;
                } else if (i == 3) {
                    for (int c = 0; c < 3; ++c) ef[c] -= T(kQM) * dr[c] / r3;
                }
            }
        }
        out[chrom] = double(ef[0] * oh[0] + ef[1] * oh[1] + ef[2] * oh[2]);
    }
}

// ---------------------------------------------------------------------------
// Site records: maps, OH unit vector, TDM position, mu, alpha
// calcIR.cu:1245-1304
// ---------------------------------------------------------------------------
template <typename T>
struct Site {
    T E, w, x, p, mup;
    T u[3], r[3], mu[3], a[6];
};

template <typename T>
void sites_t(const orc_config* cfg, const MapCoef& mc, const float* x, int natoms, const float* boxf,
             const double* eproj, std::vector<Site<T>>& s) {
    const int nam = cfg->natom_mol, ncm = cfg->nchrom_mol;
    const int nmol = natoms / nam, nchrom = nmol * ncm;
    const Box<T> box = make_box<T>(cfg, boxf);
    s.resize(nchrom);
    for (int chrom = 0; chrom < nchrom; ++chrom) {
        const int n = chrom / ncm;
        Site<T>& S = s[chrom];
        const T E = T(eproj[chrom]);
        S.E = E;
        S.w = T(mc.w[0]) + T(mc.w[1]) * E + T(mc.w[2]) * E * E;
        S.x = T(mc.x[0]) + T(mc.x[1]) * S.w;
        S.p = T(mc.p[0]) + T(mc.p[1]) * S.w;
        S.mup = T(mc.mup[0]) + T(mc.mup[1]) * E + T(mc.mup[2]) * E * E;
        const float* nof = &x[3 * (n * nam)];
        const float* nhf = &x[3 * (n * nam + h_atom(cfg, chrom))];
        T u[3];
        for (int c = 0; c < 3; ++c) u[c] = T(nhf[c]) - T(nof[c]);
        mi3<T>(u, box);
        T r = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
        for (int c = 0; c < 3; ++c) {
            u[c] /= r;
            S.u[c] = u[c];
            S.r[c] = T(nof[c]) + T(kTdmOffset) * u[c];  // calcIR.cu:1289-1291
            S.mu[c] = u[c] * S.mup * S.x;               // calcIR.cu:1294-1296
        }
        mi3<T>(S.r, box);
        const T A = T(kRamanAniso);
        S.a[0] = (A * u[0] * u[0] + T(1)) * S.x;  // calcIR.cu:1299-1304
        S.a[1] = (A * u[1] * u[1] + T(1)) * S.x;
        S.a[2] = (A * u[2] * u[2] + T(1)) * S.x;
        S.a[3] = A * u[0] * u[1] * S.x;
        S.a[4] = A * u[1] * u[2] * S.x;
        S.a[5] = A * u[2] * u[0] * S.x;
    }
}

// Hamiltonian element (n, m), calcIR.cu:1322-1401.
template <typename T>
inline T h_element(const orc_config* cfg, const MapCoef& mc, const Site<T>& Sn, const Site<T>& Sm,
                   int n, int m, const Box<T>& box) {
    const int ncm = cfg->nchrom_mol;
    if (n == m) return Sm.w - T(cfg->avef);
    const bool hod = (cfg->ispecies == 1 || cfg->ispecies == 2);  // calcIR.cu:1334, 1348
    if (hod) return T(0);
    if (n / ncm == m / ncm) {
        return (T(mc.wi[0]) + T(mc.wi[1]) * (Sn.E + Sm.E)) * Sn.x * Sm.x + T(mc.wi[2]) * Sn.p * Sm.p;
    }
    T dr[3];
    for (int c = 0; c < 3; ++c) dr[c] = Sn.r[c] - Sm.r[c];
    mi3<T>(dr, box);
    T r = std::sqrt(dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2]);
    for (int c = 0; c < 3; ++c) dr[c] /= r;
    r *= T(kBohrPerNm);
    const T uu = Sn.u[0] * Sm.u[0] + Sn.u[1] * Sm.u[1] + Sn.u[2] * Sm.u[2];
    const T ur = Sn.u[0] * dr[0] + Sn.u[1] * dr[1] + Sn.u[2] * dr[2];
    const T vr = Sm.u[0] * dr[0] + Sm.u[1] * dr[1] + Sm.u[2] * dr[2];
    return (uu - T(3) * ur * vr) / (r * r * r) * Sn.x * Sm.x * Sn.mup * Sm.mup * T(kCmPerHartree);
}

template <typename T>
void hamiltonian_t(const orc_config* cfg, const MapCoef& mc, const std::vector<Site<T>>& s,
                   const float* boxf, T* H /* row-major N*N */) {
    const int N = int(s.size());
    const Box<T> box = make_box<T>(cfg, boxf);
#pragma omp parallel for schedule(static) num_threads(threads_of(cfg))
    for (int n = 0; n < N; ++n)
        for (int m = 0; m < N; ++m) H[size_t(n) * N + m] = h_element<T>(cfg, mc, s[n], s[m], n, m, box);
}

// ---------------------------------------------------------------------------
// Symmetric eigensolver (replaces magma_ssyevd_gpu, calcIR.cu:443).
// Householder tridiagonalisation with explicit accumulation of Q, then the
// implicit-shift QL iteration on the tridiagonal matrix with the rotations
// applied to Q (Golub & Van Loan, Matrix Computations, sections 8.3.1 and
// 8.3.3).  Each QL sweep's rotations are recorded and then applied to all
// rows of Q in parallel (rows are independent).
// ---------------------------------------------------------------------------
template <typename T>
bool eigh_t(int n, std::vector<T>& A /* row-major, lower used, overwritten */, std::vector<T>& w,
            std::vector<T>& Q /* row-major out: Q[i*n+k] */, int nthreads) {
    w.assign(n, T(0));
    Q.assign(size_t(n) * n, T(0));
    for (int i = 0; i < n; ++i) Q[size_t(i) * n + i] = T(1);
    if (n == 1) {
        w[0] = A[0];
        return true;
    }
    // Symmetrise from the lower triangle (row-major n >= m).
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) A[size_t(i) * n + j] = A[size_t(j) * n + i];

    std::vector<T> d(n), e(n, T(0)), v(n), p(n), wv(n);
    for (int k = 0; k + 2 < n; ++k) {
        const int m = n - k - 1;  // length of the column below the diagonal
        // Householder vector for x = A[k+1:, k]
        T alpha = A[size_t(k + 1) * n + k];
        T sig = 0;
        for (int i = k + 2; i < n; ++i) sig += A[size_t(i) * n + k] * A[size_t(i) * n + k];
        T tau = 0, beta = alpha;
        for (int i = 0; i < n; ++i) v[i] = T(0);
        v[k + 1] = T(1);
        if (sig != T(0)) {
            T nrm = std::sqrt(alpha * alpha + sig);
            beta = (alpha >= T(0)) ? -nrm : nrm;
            tau = (beta - alpha) / beta;
            T scal = T(1) / (alpha - beta);
            for (int i = k + 2; i < n; ++i) v[i] = A[size_t(i) * n + k] * scal;
        }
        e[k] = beta;
        (void)m;
        if (tau != T(0)) {
            // p = tau * A22 v
#pragma omp parallel for schedule(static) num_threads(nthreads)
            for (int i = k + 1; i < n; ++i) {
                T acc = 0;
                const T* row = &A[size_t(i) * n];
                for (int j = k + 1; j < n; ++j) acc += row[j] * v[j];
                p[i] = tau * acc;
            }
            T pv = 0;
            for (int i = k + 1; i < n; ++i) pv += p[i] * v[i];
            const T K = T(0.5) * tau * pv;
            for (int i = k + 1; i < n; ++i) wv[i] = p[i] - K * v[i];
#pragma omp parallel for schedule(static) num_threads(nthreads)
            for (int i = k + 1; i < n; ++i) {
                T* row = &A[size_t(i) * n];
                for (int j = k + 1; j < n; ++j) row[j] -= v[i] * wv[j] + wv[i] * v[j];
            }
            // Q <- Q (I - tau v v^T)
#pragma omp parallel for schedule(static) num_threads(nthreads)
            for (int i = 0; i < n; ++i) {
                T* row = &Q[size_t(i) * n];
                T acc = 0;
                for (int j = k + 1; j < n; ++j) acc += row[j] * v[j];
                acc *= tau;
                for (int j = k + 1; j < n; ++j) row[j] -= acc * v[j];
            }
        }
    }
    for (int k = 0; k < n; ++k) d[k] = A[size_t(k) * n + k];
    e[n - 2] = A[size_t(n - 1) * n + (n - 2)];
    e[n - 1] = T(0);

    // Implicit QL with Wilkinson-type shift; e[i] couples i and i+1.
    const T eps = std::numeric_limits<T>::epsilon();
    std::vector<T> cs(n), sn(n);
    for (int l = 0; l < n; ++l) {
        int iter = 0;
        for (;;) {
            int m = l;
            for (; m < n - 1; ++m) {
                T dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
                if (std::fabs(e[m]) <= eps * dd) break;
            }
            if (m == l) break;
            if (++iter > 90) return false;
            T g = (d[l + 1] - d[l]) / (T(2) * e[l]);
            T r = std::hypot(g, T(1));
            g = d[m] - d[l] + e[l] / (g + (g >= T(0) ? r : -r));
            T s = 1, c = 1, pp = 0;
            int i = m - 1;
            bool underflow = false;
            int nrot = 0;
            for (; i >= l; --i) {
                T f = s * e[i];
                T b = c * e[i];
                r = std::hypot(f, g);
                e[i + 1] = r;
                if (r == T(0)) {
                    d[i + 1] -= pp;
                    e[m] = 0;
                    underflow = true;
                    break;
                }
                s = f / r;
                c = g / r;
                g = d[i + 1] - pp;
                r = (d[i] - g) * s + T(2) * c * b;
                pp = s * r;
                d[i + 1] = g + pp;
                g = c * r - b;
                cs[nrot] = c;
                sn[nrot] = s;
                ++nrot;
            }
            // apply the recorded rotations (columns i+1, i for i = m-1 .. ) to Q rows
            const int i0 = m - 1;
#pragma omp parallel for schedule(static) num_threads(nthreads)
            for (int row = 0; row < n; ++row) {
                T* q = &Q[size_t(row) * n];
                for (int t = 0; t < nrot; ++t) {
                    const int ii = i0 - t;
                    T f = q[ii + 1];
                    q[ii + 1] = sn[t] * q[ii] + cs[t] * f;
                    q[ii] = cs[t] * q[ii] - sn[t] * f;
                }
            }
            if (underflow) continue;
            d[l] -= pp;
            e[l] = g;
            e[m] = 0;
        }
    }
    // sort ascending
    std::vector<int> idx(n);
    for (int i = 0; i < n; ++i) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return d[a] < d[b]; });
    std::vector<T> Qs(size_t(n) * n);
    for (int k = 0; k < n; ++k) {
        w[k] = d[idx[k]];
        for (int i = 0; i < n; ++i) Qs[size_t(i) * n + k] = Q[size_t(i) * n + idx[k]];
    }
    Q.swap(Qs);
    return true;
}

// ---------------------------------------------------------------------------
// Spectral density, calcIR.cu:458-482 and 1414-1444.
// ---------------------------------------------------------------------------
template <typename T>
void spectral_density_t(const orc_config* cfg, int n, const T* w, const T* V, const T* mu /* 3*n */,
                        double* sw) {
    const int nomega = orc_nomega(cfg);
    std::vector<T> MU(3 * size_t(n), T(0));
    for (int a = 0; a < 3; ++a)
        for (int k = 0; k < n; ++k) {
            T acc = 0;
            for (int i = 0; i < n; ++i) acc += V[size_t(i) * n + k] * mu[size_t(a) * n + i];
            MU[size_t(a) * n + k] = acc;
        }
    const T gamma = T(kHbar / (double(cfg->t1) * 2.0));
    for (int i = 0; i < nomega; ++i) {
        const T om = T(float(cfg->omegaStart + float(cfg->omegaStep * i)));
        T acc = 0;
        for (int k = 0; k < n; ++k) {
            const T m2 = MU[k] * MU[k] + MU[n + k] * MU[n + k] + MU[2 * n + k] * MU[2 * n + k];
            const T dw = om - (w[k] + T(cfg->avef));
            acc += m2 * gamma / (dw * dw + gamma * gamma) / T(kPi);
        }
        sw[i] += double(acc);
    }
}

// IPR and frequency distribution, calcIR.cu:494-538.
template <typename T>
int64_t ipr_hist_t(const orc_config* cfg, int n, const T* w, const T* V, double weight, double* pw,
                   double* rw, double* rmw) {
    const int nomega = orc_nomega(cfg);
    const int ncm = cfg->nchrom_mol;
    const int nmol = n / ncm;
    int64_t skipped = 0;
    for (int k = 0; k < n; ++k) {
        T s4 = 0;
        for (int i = 0; i < n; ++i) {
            T c = V[size_t(i) * n + k];
            s4 += c * c * c * c;
        }
        const T ipr = T(1) / s4;
        T outer = 0;
        for (int mo = 0; mo < nmol; ++mo) {
            T inner = 0;
            for (int j = 0; j < ncm; ++j) {
                T c = V[size_t(mo * ncm + j) * n + k];
                inner += c * c;
            }
            outer += inner * inner;
        }
        const T mipr = T(1) / outer;
        const float wi = float(w[k]) + cfg->avef;
        const int bin = int(std::round((wi - cfg->omegaStart) / float(cfg->omegaStep)));
        if (bin < 0 || bin >= nomega) {
            // The reference warns and then writes out of bounds
            // (calcIR.cu:530-537); the bin is skipped and counted here.
            ++skipped;
            continue;
        }
        pw[bin] += weight / double(cfg->omegaStep);
        rw[bin] += weight * double(ipr) / double(cfg->omegaStep);
        rmw[bin] += weight * double(mipr) / double(cfg->omegaStep);
    }
    return skipped;
}

// ---------------------------------------------------------------------------
// Dense complex product C = A B (row-major n x n), OpenMP over rows.
// ---------------------------------------------------------------------------
template <typename T>
void cgemm_nn(int n, const std::complex<T>* A, const std::complex<T>* B, std::complex<T>* C,
              int nthreads) {
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (int i = 0; i < n; ++i) {
        std::vector<T> cr(n, T(0)), ci(n, T(0));
        for (int k = 0; k < n; ++k) {
            const T ar = A[size_t(i) * n + k].real(), ai = A[size_t(i) * n + k].imag();
            const std::complex<T>* brow = &B[size_t(k) * n];
            for (int j = 0; j < n; ++j) {
                const T br = brow[j].real(), bi = brow[j].imag();
                cr[j] += ar * br - ai * bi;
                ci[j] += ar * bi + ai * br;
            }
        }
        for (int j = 0; j < n; ++j) C[size_t(i) * n + j] = std::complex<T>(cr[j], ci[j]);
    }
}

// C = A B^T (plain transpose), row-major.
template <typename T>
void cgemm_nt(int n, const std::complex<T>* A, const std::complex<T>* B, std::complex<T>* C,
              int nthreads) {
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < n; ++j) {
            T cr = 0, ci = 0;
            const std::complex<T>* a = &A[size_t(i) * n];
            const std::complex<T>* b = &B[size_t(j) * n];
            for (int k = 0; k < n; ++k) {
                cr += a[k].real() * b[k].real() - a[k].imag() * b[k].imag();
                ci += a[k].real() * b[k].imag() + a[k].imag() * b[k].real();
            }
            C[size_t(i) * n + j] = std::complex<T>(cr, ci);
        }
    }
}

// The nine per-chromophore vectors of the correlation functions, in the
// order mu_x, mu_y, mu_z, a_xx, a_yy, a_zz, a_xy, a_yz, a_zx.
template <typename T>
void collect_m(const std::vector<Site<T>>& s, std::vector<T>& M /* 9*n */) {
    const int n = int(s.size());
    M.resize(9 * size_t(n));
    for (int i = 0; i < n; ++i) {
        for (int c = 0; c < 3; ++c) M[size_t(c) * n + i] = s[i].mu[c];
        for (int c = 0; c < 6; ++c) M[size_t(3 + c) * n + i] = s[i].a[c];
    }
}

// (b, a) pairs of calcIR.cu:628-733: value = M_b(t)^T F M_a(0).
struct Pair {
    int b, a;
};
constexpr Pair kIR[3] = {{0, 0}, {1, 1}, {2, 2}};
constexpr Pair kIIFII[3] = {{3, 3}, {4, 4}, {5, 5}};
constexpr Pair kIJFIJ[3] = {{6, 6}, {7, 7}, {8, 8}};
constexpr Pair kIIFJJ[6] = {{4, 3}, {5, 3}, {3, 4}, {5, 4}, {3, 5}, {4, 5}};

template <typename T>
double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

template <typename T>
int run_t(const orc_config* cfg, int natoms, int nframes, const float* xyz, const float* box,
          const float* times, orc_result* out) {
    using C = std::complex<T>;
    MapCoef mc;
    if (!cfg_coef(cfg, &mc)) return 2;
    if (!layout_ok(cfg, natoms)) return 1;
    if (nframes < 2 || cfg->ntcfpoints <= 0 || cfg->nsamples <= 0) return 1;
    const int nthreads = threads_of(cfg);
    const int nmol = natoms / cfg->natom_mol;
    const int N = nmol * cfg->nchrom_mol;
    const int ntcf = cfg->ntcfpoints;
    const int nomega = orc_nomega(cfg);
    const int nfft = ntcf + cfg->nzeros;

    // calcIR.cu:342-345 with the XTC precision of 1000
    // calcIR.cu:342-345: the first spacing rounded at the trajectory's precision
    const float tprec = cfg->time_prec > 0.0f ? cfg->time_prec : 1000.0f;
    const float frame_dt = float(std::round((times[1] - times[0]) * tprec) / (1. * tprec));

    std::vector<C> tcf(ntcf), tcfvv(ntcf), tcfvh(ntcf);
    std::vector<double> sw(nomega, 0.0), pw(nomega, 0.0), rw(nomega, 0.0), rmw(nomega, 0.0);
    std::vector<double> eproj(N);
    std::vector<Site<T>> s;
    std::vector<T> Hm(size_t(N) * N), w, V, M, M0;
    std::vector<C> F, Y, U, tmp, P;
    const bool literal = cfg->propagation == ORC_PROP_LITERAL;
    if (literal) {
        F.assign(size_t(N) * N, C(0));
        U.assign(size_t(N) * N, C(0));
        tmp.assign(size_t(N) * N, C(0));
        P.assign(size_t(N) * N, C(0));
    }
    std::vector<C> Vc;
    std::vector<double> fsw(nomega);
    out->nsteps = 0;
    out->hist_skipped = 0;
    out->t_hamiltonian = out->t_eigh = out->t_hist = out->t_propagate = out->t_finalize = 0;

    for (int smp = 0; smp < cfg->nsamples; ++smp) {
        for (int k = 0; k < ntcf; ++k) {
            float desired;
            const int f = orc_frame_index(cfg, smp, k, frame_dt, &desired);
            if (f < 0 || f >= nframes) return 3;
            if (std::fabs(desired - times[f]) > frame_dt * 1E-1) return 3;  // calcIR.cu:391
            const float* x = &xyz[size_t(f) * natoms * 3];
            const float* bx = &box[size_t(f) * box_floats(cfg)];

            double t0 = now_s<T>();
            eproj_t<T>(cfg, x, natoms, bx, eproj.data());
            sites_t<T>(cfg, mc, x, natoms, bx, eproj.data(), s);
            hamiltonian_t<T>(cfg, mc, s, bx, Hm.data());
            double t1 = now_s<T>();
            std::vector<T> A(Hm);
            if (!eigh_t<T>(N, A, w, V, nthreads)) return 4;
            double t2 = now_s<T>();
            collect_m<T>(s, M);
            if (k == 0) {
                std::fill(fsw.begin(), fsw.end(), 0.0);
                spectral_density_t<T>(cfg, N, w.data(), V.data(), M.data(), fsw.data());
                for (int i = 0; i < nomega; ++i) sw[i] += double(T(fsw[i]));
            }
            out->hist_skipped += ipr_hist_t<T>(cfg, N, w.data(), V.data(), 1.0, pw.data(), rw.data(),
                                               rmw.data());
            double t3 = now_s<T>();

            // calcIR.cu:577-617 (literal) or the algebraically identical
            // vector form Y = F M(0).
            if (k == 0) {
                M0 = M;
                if (literal) {
                    std::fill(F.begin(), F.end(), C(0));
                    for (int i = 0; i < N; ++i) F[size_t(i) * N + i] = C(1);
                } else {
                    Y.assign(9 * size_t(N), C(0));
                    for (size_t i = 0; i < Y.size(); ++i) Y[i] = C(M0[i]);
                }
            } else {
                const T hbar = T(kHbar);
                std::vector<C> ph(N);
                for (int i = 0; i < N; ++i) {
                    // calcIR.cu:1512-1514, P = exp(+i w dt / hbar)
                    T arg = w[i] * T(cfg->dt) / hbar;
                    ph[i] = C(std::cos(arg), std::sin(arg));
                }
                if (literal) {
                    Vc.assign(size_t(N) * N, C(0));
                    for (size_t i = 0; i < Vc.size(); ++i) Vc[i] = C(V[i]);
                    std::fill(P.begin(), P.end(), C(0));
                    for (int i = 0; i < N; ++i) P[size_t(i) * N + i] = ph[i];
                    cgemm_nn<T>(N, Vc.data(), P.data(), tmp.data(), nthreads);   // V P
                    cgemm_nt<T>(N, tmp.data(), Vc.data(), U.data(), nthreads);   // (V P) V^T
                    cgemm_nn<T>(N, U.data(), F.data(), tmp.data(), nthreads);    // U F
                    F.swap(tmp);
                } else {
                    // Y <- V diag(ph) V^T Y for the 9 columns
                    std::vector<C> G(9 * size_t(N));
#pragma omp parallel for schedule(static) num_threads(nthreads)
                    for (int kk = 0; kk < N; ++kk) {
                        for (int a = 0; a < 9; ++a) {
                            C acc(0);
                            for (int i = 0; i < N; ++i) acc += V[size_t(i) * N + kk] * Y[size_t(a) * N + i];
                            G[size_t(a) * N + kk] = acc * ph[kk];
                        }
                    }
#pragma omp parallel for schedule(static) num_threads(nthreads)
                    for (int i = 0; i < N; ++i) {
                        const T* vrow = &V[size_t(i) * N];
                        for (int a = 0; a < 9; ++a) {
                            C acc(0);
                            for (int kk = 0; kk < N; ++kk) acc += vrow[kk] * G[size_t(a) * N + kk];
                            Y[size_t(a) * N + i] = acc;
                        }
                    }
                }
            }
            // bilinear forms M_b(t)^T F M_a(0)
            auto form = [&](int b, int a) -> C {
                C acc(0);
                if (literal) {
                    for (int i = 0; i < N; ++i) {
                        C fy(0);
                        const C* frow = &F[size_t(i) * N];
                        for (int j = 0; j < N; ++j) fy += frow[j] * M0[size_t(a) * N + j];
                        acc += M[size_t(b) * N + i] * fy;
                    }
                } else {
                    for (int i = 0; i < N; ++i) acc += M[size_t(b) * N + i] * Y[size_t(a) * N + i];
                }
                return acc;
            };
            C ir(0), iiii(0), ijij(0), iijj(0);
            for (const Pair& q : kIR) ir += form(q.b, q.a);
            for (const Pair& q : kIIFII) iiii += form(q.b, q.a);
            for (const Pair& q : kIJFIJ) ijij += form(q.b, q.a);
            for (const Pair& q : kIIFJJ) iijj += form(q.b, q.a);
            // calcIR.cu:646-648, 736-745
            tcf[k] += ir;
            tcfvv[k] += (T(3) * iiii + iijj + T(4) * ijij) / T(15);
            tcfvh[k] += (T(2) * iiii - iijj + T(6) * ijij) / T(30);
            double t4 = now_s<T>();
            out->t_hamiltonian += t1 - t0;
            out->t_eigh += t2 - t1;
            out->t_hist += t3 - t2;
            out->t_propagate += t4 - t3;
            out->nsteps += 1;
        }
    }

    double t5 = now_s<T>();
    // calcIR.cu:775-827: decay, 1/nsamples, zero pad, C2R
    std::vector<double> X(2 * size_t(ntcf));
    auto finish = [&](std::vector<C>& tc, double* tcf_out, double* ft_out) {
        for (int i = 0; i < ntcf; ++i) {
            const T dcy = T(std::exp(-1.0 * i * cfg->dt / (2.0 * cfg->t1)) / (1. * cfg->nsamples));
            tc[i] = tc[i] * dcy;
            X[2 * i] = double(tc[i].real());
            X[2 * i + 1] = double(tc[i].imag());
            if (tcf_out) {
                tcf_out[2 * i] = X[2 * i];
                tcf_out[2 * i + 1] = X[2 * i + 1];
            }
        }
        if (ft_out) orc_c2r(nfft, ntcf, X.data(), ft_out);
    };
    finish(tcf, out->tcf_ir, out->ftcf_ir);
    finish(tcfvv, out->tcf_vv, out->ftcf_vv);
    finish(tcfvh, out->tcf_vh, out->ftcf_vh);
    // calcIR.cu:831-839
    const double norm = double(N) * cfg->nsamples * ntcf;
    for (int i = 0; i < nomega; ++i) {
        if (out->sw) out->sw[i] = sw[i] / double(cfg->nsamples);
        const double p = pw[i] / norm;
        if (out->pw) out->pw[i] = p;
        if (out->rw) out->rw[i] = (rw[i] / norm) / p;
        if (out->rmw) out->rmw[i] = (rmw[i] / norm) / p;
    }
    out->t_finalize = now_s<T>() - t5;
    return 0;
}

template <typename T>
int eigh_entry(int n, const double* H, double* w, double* V, int nthreads) {
    std::vector<T> A(size_t(n) * n), wv, Q;
    for (size_t i = 0; i < A.size(); ++i) A[i] = T(H[i]);
    if (!eigh_t<T>(n, A, wv, Q, nthreads)) return 4;
    for (int i = 0; i < n; ++i) w[i] = double(wv[i]);
    for (size_t i = 0; i < Q.size(); ++i) V[i] = double(Q[i]);
    return 0;
}

template <typename T>
void export_sites(const std::vector<Site<T>>& s, double* out) {
    for (size_t i = 0; i < s.size(); ++i) {
        double* o = out + i * ORC_SITE_STRIDE;
        const Site<T>& S = s[i];
        o[ORC_SITE_E] = S.E;
        o[ORC_SITE_W] = S.w;
        o[ORC_SITE_X] = S.x;
        o[ORC_SITE_P] = S.p;
        o[ORC_SITE_MUP] = S.mup;
        for (int c = 0; c < 3; ++c) {
            o[ORC_SITE_U + c] = S.u[c];
            o[ORC_SITE_R + c] = S.r[c];
            o[ORC_SITE_MU + c] = S.mu[c];
        }
        for (int c = 0; c < 6; ++c) o[ORC_SITE_A + c] = S.a[c];
        for (int c = ORC_SITE_A + 6; c < ORC_SITE_STRIDE; ++c) o[c] = 0;
    }
}

}  // namespace

extern "C" {

double orc_minimage(int prec, double dx, double boxl) {
    if (prec == ORC_F32) return double(minimage<float>(float(dx), float(boxl)));
    return minimage<double>(dx, boxl);
}

int orc_nomega(const orc_config* cfg) {
    // calcIR.cu:94, float arithmetic then truncation
    return int((cfg->omegaStop - cfg->omegaStart) / float(cfg->omegaStep) + 1);
}

int orc_frame_index(const orc_config* cfg, int sample, int point, float frame_dt, float* desired_time) {
    // calcIR.cu:378-379, float arithmetic as in the reference
    float t = float(sample) * cfg->sampleEvery + cfg->beginTime + cfg->dt * float(point);
    if (desired_time) *desired_time = t;
    return int(std::round(t / frame_dt));
}

int orc_eproj(int prec, const orc_config* cfg, const float* x, int natoms, const float box[3],
              double* eproj) {
    if (!cfg || !x || !eproj || !layout_ok(cfg, natoms)) return 1;
    if (prec == ORC_F32)
        eproj_t<float>(cfg, x, natoms, box, eproj);
    else
        eproj_t<double>(cfg, x, natoms, box, eproj);
    return 0;
}

int orc_sites(int prec, const orc_config* cfg, const float* x, int natoms, const float box[3],
              const double* eproj, double* site) {
    MapCoef mc;
    if (!cfg || !cfg_coef(cfg, &mc)) return 2;
    if (!x || !eproj || !site || !layout_ok(cfg, natoms)) return 1;
    if (prec == ORC_F32) {
        std::vector<Site<float>> s;
        sites_t<float>(cfg, mc, x, natoms, box, eproj, s);
        export_sites(s, site);
    } else {
        std::vector<Site<double>> s;
        sites_t<double>(cfg, mc, x, natoms, box, eproj, s);
        export_sites(s, site);
    }
    return 0;
}

int orc_hamiltonian(int prec, const orc_config* cfg, const float* x, int natoms, const float box[3],
                    const double* eproj, double* H) {
    MapCoef mc;
    if (!cfg || !cfg_coef(cfg, &mc)) return 2;
    if (!x || !eproj || !H || !layout_ok(cfg, natoms)) return 1;
    const int N = (natoms / cfg->natom_mol) * cfg->nchrom_mol;
    if (prec == ORC_F32) {
        std::vector<Site<float>> s;
        sites_t<float>(cfg, mc, x, natoms, box, eproj, s);
        std::vector<float> h(size_t(N) * N);
        hamiltonian_t<float>(cfg, mc, s, box, h.data());
        for (size_t i = 0; i < h.size(); ++i) H[i] = h[i];
    } else {
        std::vector<Site<double>> s;
        sites_t<double>(cfg, mc, x, natoms, box, eproj, s);
        hamiltonian_t<double>(cfg, mc, s, box, H);
    }
    return 0;
}

int orc_eigh(int prec, int n, const double* H, double* w, double* V, int nthreads) {
    if (n <= 0 || !H || !w || !V) return 1;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#else
    nthreads = 1;
#endif
    return prec == ORC_F32 ? eigh_entry<float>(n, H, w, V, nthreads)
                           : eigh_entry<double>(n, H, w, V, nthreads);
}

int orc_spectral_density(int prec, const orc_config* cfg, int n, const double* w, const double* V,
                         const double* mu, double* sw) {
    if (prec == ORC_F32) {
        std::vector<float> wf(w, w + n), Vf(V, V + size_t(n) * n), mf(mu, mu + 3 * size_t(n));
        spectral_density_t<float>(cfg, n, wf.data(), Vf.data(), mf.data(), sw);
    } else {
        spectral_density_t<double>(cfg, n, w, V, mu, sw);
    }
    return 0;
}

int64_t orc_ipr_hist(int prec, const orc_config* cfg, int n, const double* w, const double* V,
                     double weight, double* pw, double* rw, double* rmw) {
    if (prec == ORC_F32) {
        std::vector<float> wf(w, w + n), Vf(V, V + size_t(n) * n);
        return ipr_hist_t<float>(cfg, n, wf.data(), Vf.data(), weight, pw, rw, rmw);
    }
    return ipr_hist_t<double>(cfg, n, w, V, weight, pw, rw, rmw);
}

int orc_c2r(int n, int nx, const double* X, double* y) {
    if (n <= 0 || nx < 0) return 1;
    const int half = n / 2;  // inputs X[0..half] are read
    const int jmax = std::min(nx - 1, half);
    for (int k = 0; k < n; ++k) {
        double acc = nx > 0 ? X[0] : 0.0;  // Im X[0] ignored
        for (int j = 1; j <= jmax; ++j) {
            const long long jk = (static_cast<long long>(j) * k) % n;
            const double ang = kTwoPiExact * double(jk) / double(n);
            const double re = X[2 * j], im = X[2 * j + 1];
            if ((n % 2 == 0) && j == half) {
                acc += re * std::cos(ang);  // X[n/2] taken as real
            } else {
                acc += 2.0 * (re * std::cos(ang) - im * std::sin(ang));
            }
        }
        y[k] = acc;
    }
    return 0;
}

int orc_run(int prec, const orc_config* cfg, int natoms, int nframes, const float* xyz,
            const float* box, const float* times, orc_result* out) {
    if (!cfg || !xyz || !box || !times || !out) return 1;
    return prec == ORC_F32 ? run_t<float>(cfg, natoms, nframes, xyz, box, times, out)
                           : run_t<double>(cfg, natoms, nframes, xyz, box, times, out);
}

}  // extern "C"
