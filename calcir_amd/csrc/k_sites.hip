// k_sites.hip -- electric-field projection fused with the site records, and
// the exciton Hamiltonian build (gfx950).
//
// Field + sites (calcIR.cu:1020-1168 and the per-row part of get_kappa_GPU,
// calcIR.cu:1245-1304): one lane per chromophore; the molecules of a frame
// are staged in LDS tiles as SoA (FP32 oxygen for the cutoff decision,
// real_t H1/H2/M' with the M site already relocated), so every lane reads
// the tile from LDS instead of the strided AoS global array.  Field sums are
// Neumaier-compensated.
//
// Hamiltonian (calcIR.cu:1306-1404): a 64x64 tile per workgroup over the
// site records; every element is computed once, the tile is written
// column-major with lanes along the row index (coalesced), and the
// transition-dipole formula is evaluated in a symmetric operation order so
// that H(i,j) == H(j,i) bit for bit.

#include "common.hpp"
#include "sites.hpp"
#include "spectra.hpp"

namespace cir {

constexpr int kFieldBlock = 256;
constexpr int kFieldTile = 256;  // molecules staged per LDS tile

// The discrete cutoff decision (calcIR.cu:1106-1113), always in FP32 with
// the reference's operation order, identical in every precision form.
__device__ __forceinline__ bool within_cutoff(float ox, float oy, float oz, float hx, float hy,
                                              float hz, const PBox<float>& B, float cutoff) {
#pragma clang fp contract(off)
    float d0 = ox - hx, d1 = oy - hy, d2 = oz - hz;
    minimage3<float>(d0, d1, d2, B);
    const float r = __fsqrt_rn(d0 * d0 + d1 * d1 + d2 * d2);
    return !(r > cutoff);
}

template <typename T>
__global__ void __launch_bounds__(kFieldBlock)
eproj_sites_kernel(int natoms, int natom_mol, int nchrom_mol, int imodel, int rules, float cutoff,
                   const float* __restrict__ xyz, const float* __restrict__ box, int boxk, MapT<T> mc,
                   T* __restrict__ sites) {
    __shared__ float sO[3][kFieldTile];
    __shared__ T sA[9][kFieldTile];  // H1, H2, negative site (x,y,z each)

    const int f = blockIdx.y;
    const int nmol = natoms / natom_mol;
    const int N = nmol * nchrom_mol;
    const float* x = xyz + size_t(f) * natoms * 3;
    const PBox<float> Bf = load_box<float>(box, boxk, f);
    const PBox<T> Bt = load_box<T>(box, boxk, f);
    // calcIR.cu:1117-1160: the charged sites of molecule m are its atoms
    // 1..natom_mol-1, +0.52 on atoms 1 and 2 and -1.04 on atom 3 (later atoms
    // carry nothing), so three-site water has no negative site.  The
    // msite_bisector rule builds one for it instead.
    const bool built_m = (rules & kRuleMsiteBisector) && natom_mol == 3;
    const bool have_m = natom_mol >= 4 || built_m;

    const int chrom = blockIdx.x * kFieldBlock + threadIdx.x;
    const bool active = chrom < N;
    const int n = active ? chrom / nchrom_mol : 0;
    // calcIR.cu:1061-1070 picks H1 or H2 by chrom % 2 whatever nchrom_mol is
    // (with nchrom_mol 1, even molecules use H1 and odd ones H2); the
    // oh_per_molecule rule takes H(1 + chrom % nchrom_mol).
    const int hatom = 1 + ((rules & kRuleOhPerMolecule) ? chrom % nchrom_mol : chrom % 2);
    float hxf = 0, hyf = 0, hzf = 0;
    if (active) {
        hxf = x[3 * (n * natom_mol + hatom) + 0];
        hyf = x[3 * (n * natom_mol + hatom) + 1];
        hzf = x[3 * (n * natom_mol + hatom) + 2];
    }
    const T hx = T(hxf), hy = T(hyf), hz = T(hzf);
    const T bohr = T(kBohrPerNm);
    Compensated<T> ex, ey, ez;

    for (int m0 = 0; m0 < nmol; m0 += kFieldTile) {
        const int cnt = min(kFieldTile, nmol - m0);
        __syncthreads();
        for (int t = threadIdx.x; t < cnt; t += kFieldBlock) {
            const float* mx = x + 3 * size_t((m0 + t) * natom_mol);
            const float ox = mx[0], oy = mx[1], oz = mx[2];
            sO[0][t] = ox;
            sO[1][t] = oy;
            sO[2][t] = oz;
            for (int c = 0; c < 3; ++c) {
                sA[0 + c][t] = T(mx[3 + c]);
                sA[3 + c][t] = T(mx[6 + c]);
            }
            const T o[3] = {T(ox), T(oy), T(oz)};
            T m[3] = {T(0), T(0), T(0)};
            if (built_m) {
                // the TIP4P negative site on the HOH bisector, 0.015 nm from O
                T h1[3] = {T(mx[3]) - o[0], T(mx[4]) - o[1], T(mx[5]) - o[2]};
                T h2[3] = {T(mx[6]) - o[0], T(mx[7]) - o[1], T(mx[8]) - o[2]};
                minimage3<T>(h1[0], h1[1], h1[2], Bt);
                minimage3<T>(h2[0], h2[1], h2[2], Bt);
                const T r1 = sqrt(h1[0] * h1[0] + h1[1] * h1[1] + h1[2] * h1[2]);
                const T r2 = sqrt(h2[0] * h2[0] + h2[1] * h2[1] + h2[2] * h2[2]);
                T bi[3];
                for (int c = 0; c < 3; ++c) bi[c] = h1[c] / r1 + h2[c] / r2;
                const T rb = sqrt(bi[0] * bi[0] + bi[1] * bi[1] + bi[2] * bi[2]);
                for (int c = 0; c < 3; ++c) m[c] = o[c] + T(kOMTip4p) * bi[c] / rb;
            } else if (have_m) {
                m[0] = T(mx[9]);
                m[1] = T(mx[10]);
                m[2] = T(mx[11]);
                if (imodel != 0) {
                    // calcIR.cu:1124-1140: M moved to 0.015 nm along O->M
                    T om[3] = {m[0] - o[0], m[1] - o[1], m[2] - o[2]};
                    minimage3<T>(om[0], om[1], om[2], Bt);
                    const T rom = sqrt(om[0] * om[0] + om[1] * om[1] + om[2] * om[2]);
                    for (int c = 0; c < 3; ++c) m[c] = o[c] + T(kOMTip4p) * om[c] / rom;
                }
            }
            for (int c = 0; c < 3; ++c) sA[6 + c][t] = m[c];
        }
        __syncthreads();
        if (!active) continue;
        for (int t = 0; t < cnt; ++t) {
            const int m = m0 + t;
            if (m == n) continue;  // calcIR.cu:1099
            if (!within_cutoff(sO[0][t], sO[1][t], sO[2][t], hxf, hyf, hzf, Bf, cutoff)) continue;
            for (int s = 0; s < (have_m ? 3 : 2); ++s) {
                T dx = hx - sA[3 * s + 0][t], dy = hy - sA[3 * s + 1][t], dz = hz - sA[3 * s + 2][t];
                minimage3<T>(dx, dy, dz, Bt);
                dx = dx * bohr;
                dy = dy * bohr;
                dz = dz * bohr;
                const T r = sqrt(dx * dx + dy * dy + dz * dz);
                const T q = (s < 2) ? T(kQH) : T(-kQM);
                const T qr3 = q / (r * r * r);
                ex.add(qr3 * dx);
                ey.add(qr3 * dy);
                ez.add(qr3 * dz);
            }
        }
    }
    if (!active) return;

    // own OH unit vector, TDM position and maps (calcIR.cu:1047-1080, 1253-1304)
    const float* no = x + 3 * size_t(n * natom_mol);
    const T ox = T(no[0]), oy = T(no[1]), oz = T(no[2]);
    T ux = hx - ox, uy = hy - oy, uz = hz - oz;
    minimage3<T>(ux, uy, uz, Bt);
    const T r = sqrt(ux * ux + uy * uy + uz * uz);
    ux /= r;
    uy /= r;
    uz /= r;
    const T E = ex.value() * ux + ey.value() * uy + ez.value() * uz;
    const T w = mc.w0 + mc.w1 * E + mc.w2 * E * E;
    const T xx = mc.x0 + mc.x1 * w;
    const T p = mc.p0 + mc.p1 * w;
    const T mup = mc.m0 + mc.m1 * E + mc.m2 * E * E;
    const T A = T(kRamanAniso);
    T* S = sites + size_t(f) * CIR_NSITE * N;
    S[CIR_SITE_E * N + chrom] = E;
    S[CIR_SITE_W * N + chrom] = w;
    S[CIR_SITE_X * N + chrom] = xx;
    S[CIR_SITE_P * N + chrom] = p;
    S[CIR_SITE_MUP * N + chrom] = mup;
    S[(CIR_SITE_U + 0) * N + chrom] = ux;
    S[(CIR_SITE_U + 1) * N + chrom] = uy;
    S[(CIR_SITE_U + 2) * N + chrom] = uz;
    T rx = ox + T(kTdmOffset) * ux, ry = oy + T(kTdmOffset) * uy, rz = oz + T(kTdmOffset) * uz;
    minimage3<T>(rx, ry, rz, Bt);  // calcIR.cu:1289-1291 (wrapped; only differences are used)
    S[(CIR_SITE_R + 0) * N + chrom] = rx;
    S[(CIR_SITE_R + 1) * N + chrom] = ry;
    S[(CIR_SITE_R + 2) * N + chrom] = rz;
    S[(CIR_SITE_MU + 0) * N + chrom] = ux * mup * xx;
    S[(CIR_SITE_MU + 1) * N + chrom] = uy * mup * xx;
    S[(CIR_SITE_MU + 2) * N + chrom] = uz * mup * xx;
    S[(CIR_SITE_A + 0) * N + chrom] = (A * ux * ux + T(1)) * xx;
    S[(CIR_SITE_A + 1) * N + chrom] = (A * uy * uy + T(1)) * xx;
    S[(CIR_SITE_A + 2) * N + chrom] = (A * uz * uz + T(1)) * xx;
    S[(CIR_SITE_A + 3) * N + chrom] = A * ux * uy * xx;
    S[(CIR_SITE_A + 4) * N + chrom] = A * uy * uz * xx;
    S[(CIR_SITE_A + 5) * N + chrom] = A * uz * ux * xx;
}

constexpr int kHTile = 64;

template <typename T>
__global__ void __launch_bounds__(256)
hamiltonian_kernel(int N, int ld, int nchrom_mol, int hod, T avef, T wi0, T wi1, T wi2,
                   const float* __restrict__ box, int boxk, const T* __restrict__ sites, T* __restrict__ H) {
    // per-tile site data: E, w, x, p, s = x*mup, u(3), r(3)
    __shared__ T rs[11][kHTile];
    __shared__ T cs[11][kHTile];
    const int f = blockIdx.z;
    const int i0 = blockIdx.x * kHTile, j0 = blockIdx.y * kHTile;
    const T* S = sites + size_t(f) * CIR_NSITE * N;
    const int tid = threadIdx.x;
    if (tid < 2 * kHTile) {
        const bool row = tid < kHTile;
        const int l = row ? tid : tid - kHTile;
        const int g = (row ? i0 : j0) + l;
        T(*d)[kHTile] = row ? rs : cs;
        if (g < N) {
            d[0][l] = S[CIR_SITE_E * N + g];
            d[1][l] = S[CIR_SITE_W * N + g];
            d[2][l] = S[CIR_SITE_X * N + g];
            d[3][l] = S[CIR_SITE_P * N + g];
            d[4][l] = S[CIR_SITE_X * N + g] * S[CIR_SITE_MUP * N + g];
            for (int c = 0; c < 3; ++c) {
                d[5 + c][l] = S[(CIR_SITE_U + c) * N + g];
                d[8 + c][l] = S[(CIR_SITE_R + c) * N + g];
            }
        } else {
            for (int c = 0; c < 11; ++c) d[c][l] = T(0);
        }
    }
    __syncthreads();
    const PBox<T> bb = load_box<T>(box, boxk, f);
    const int tx = tid % kHTile, ty = tid / kHTile;
    const int i = i0 + tx;
    if (i >= ld) return;
    T* Hf = H + size_t(f) * N * ld;
    for (int q = 0; q < kHTile / 4; ++q) {
        const int jl = ty + 4 * q;
        const int j = j0 + jl;
        if (j >= N) break;
        T h;
        if (i >= N) {
            h = T(0);  // padding rows of the leading dimension
        } else if (i == j) {
            h = rs[1][tx] - avef;  // calcIR.cu:1322-1327
        } else if (hod) {
            h = T(0);  // calcIR.cu:1334-1336, 1348-1350
        } else if (i / nchrom_mol == j / nchrom_mol) {
            // calcIR.cu:1330-1341
            h = (wi0 + wi1 * (rs[0][tx] + cs[0][jl])) * (rs[2][tx] * cs[2][jl]) +
                wi2 * (rs[3][tx] * cs[3][jl]);
        } else {
            // calcIR.cu:1353-1401, symmetric operation order
            const T g = tdc_geometry<T>(rs[5][tx], rs[6][tx], rs[7][tx], rs[8][tx], rs[9][tx], rs[10][tx],
                                        cs[5][jl], cs[6][jl], cs[7][jl], cs[8][jl], cs[9][jl], cs[10][jl], bb);
            h = g * (rs[4][tx] * cs[4][jl]) * T(kCmPerHartree);
        }
        Hf[size_t(j) * ld + i] = h;
    }
}

template <typename T>
int launch_eproj_sites(int nframes, int natoms, int natom_mol, int nchrom_mol, int imodel, int rules,
                       float cutoff, const float* xyz, const float* box, int boxk, const MapCoef& mc,
                       void* sites, hipStream_t s) {
    const int N = (natoms / natom_mol) * nchrom_mol;
    dim3 grid(unsigned(ceil_div(N, kFieldBlock)), unsigned(nframes));
    hipLaunchKernelGGL(eproj_sites_kernel<T>, grid, dim3(kFieldBlock), 0, s, natoms, natom_mol,
                       nchrom_mol, imodel, rules, cutoff, xyz, box, boxk, map_t<T>(mc), static_cast<T*>(sites));
    CIR_CHECK_LAUNCH();
    return CIR_OK;
}

template <typename T>
int launch_hamiltonian(int nframes, int N, int ld, int nchrom_mol, int ispecies, const MapCoef& mc,
                       float avef, const float* box, int boxk, const void* sites, void* H, hipStream_t s) {
    dim3 grid(unsigned(ceil_div(ld, kHTile)), unsigned(ceil_div(N, kHTile)), unsigned(nframes));
    const int hod = (ispecies == 1 || ispecies == 2) ? 1 : 0;
    hipLaunchKernelGGL(hamiltonian_kernel<T>, grid, dim3(256), 0, s, N, ld, nchrom_mol, hod, T(avef),
                       T(mc.wi[0]), T(mc.wi[1]), T(mc.wi[2]), box, boxk, static_cast<const T*>(sites),
                       static_cast<T*>(H));
    CIR_CHECK_LAUNCH();
    return CIR_OK;
}

int eproj_sites(int prec, int nframes, int natoms, int natom_mol, int nchrom_mol, int imodel, int rules,
                const MapCoef& mc, float cutoff, const float* xyz_d, const float* box_d, int boxk, void* sites_d,
                hipStream_t s) {
    // atoms 1 and 2 of a molecule are its hydrogens (calcIR.cu:1061-1070)
    if (nframes <= 0 || natom_mol < 3 || nchrom_mol < 1 || natoms <= 0 || natoms % natom_mol || !xyz_d ||
        !box_d || !sites_d || (boxk != 0 && boxk != 1) || (rules & ~(kRuleOhPerMolecule | kRuleMsiteBisector)))
        return fail(CIR_ERR_ARG, "cir_eproj_sites: bad sizes or null pointer");
    if ((rules & kRuleOhPerMolecule) && nchrom_mol > 2)
        return fail(CIR_ERR_ARG, "cir_eproj_sites: one or two OH stretches per molecule with oh_per_molecule");
    if (int64_t(nframes) > 65535) return fail(CIR_ERR_ARG, "cir_eproj_sites: nframes > 65535");
    if (int64_t(natoms / natom_mol) * nchrom_mol > (int64_t(1) << 30))
        return fail(CIR_ERR_ARG, "cir_eproj_sites: too many chromophores");
    if (prec == CIR_PREC_FP64)
        return launch_eproj_sites<double>(nframes, natoms, natom_mol, nchrom_mol, imodel, rules, cutoff, xyz_d,
                                          box_d, boxk, mc, sites_d, s);
    return launch_eproj_sites<float>(nframes, natoms, natom_mol, nchrom_mol, imodel, rules, cutoff, xyz_d, box_d,
                                     boxk, mc, sites_d, s);
}

int hamiltonian(int prec, int nframes, int n, int ld, int nchrom_mol, int ispecies, const MapCoef& mc, float avef,
                const float* box_d, int boxk, const void* sites_d, void* H_d, hipStream_t s) {
    if (nframes <= 0 || n <= 0 || ld < n || nchrom_mol <= 0 || !box_d || !sites_d || !H_d ||
        (boxk != 0 && boxk != 1))
        return fail(CIR_ERR_ARG, "cir_hamiltonian: bad sizes or null pointer");
    if (nframes > 65535) return fail(CIR_ERR_ARG, "cir_hamiltonian: nframes > 65535");
    if (prec == CIR_PREC_FP64)
        return launch_hamiltonian<double>(nframes, n, ld, nchrom_mol, ispecies, mc, avef, box_d, boxk, sites_d, H_d,
                                          s);
    return launch_hamiltonian<float>(nframes, n, ld, nchrom_mol, ispecies, mc, avef, box_d, boxk, sites_d, H_d, s);
}

}  // namespace cir

using namespace cir;

extern "C" int cir_eproj_sites(int prec, int nframes, int natoms, int natom_mol, int nchrom_mol,
                               int imodel, int ispecies, int imap, float cutoff,
                               const float* xyz_d, const float* box_d, void* sites_d,
                               void* stream) {
    MapCoef mc;
    if (!map_coef(imap, ispecies, &mc)) return fail(CIR_ERR_MAP, "map/species not in the map table");
    return eproj_sites(prec, nframes, natoms, natom_mol, nchrom_mol, imodel, 0, mc, cutoff, xyz_d, box_d, 0,
                       sites_d, as_stream(stream));
}

extern "C" int cir_hamiltonian(int prec, int nframes, int n, int ld, int nchrom_mol, int ispecies,
                               int imap, float avef, const float* box_d, const void* sites_d,
                               void* H_d, void* stream) {
    MapCoef mc;
    if (!map_coef(imap, ispecies, &mc)) return fail(CIR_ERR_MAP, "map/species not in the map table");
    return hamiltonian(prec, nframes, n, ld, nchrom_mol, ispecies, mc, avef, box_d, 0, sites_d, H_d,
                       as_stream(stream));
}
