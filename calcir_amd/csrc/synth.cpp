// synth.cpp -- seeded synthetic rigid-water trajectories.
//
// The reference reads GROMACS XTC trajectories (calcIR.cu:212-227, 341-398).
// This source produces frames of the same layout (natoms x rvec, atom order
// O, H1, H2, M per molecule, nm, cubic box, t = k*0.010 ps from 0) from a
// seeded Ornstein-Uhlenbeck rigid-body model, deterministic across machines:
//   * oxygen sites on the smallest simple-cubic lattice with >= n sites,
//     occupied by a seeded shuffle, uniform jitter of +-0.02 nm;
//   * uniform random orientations (Shoemake quaternions), redrawn while an
//     intermolecular H...H distance is below 0.18 nm or an H...M distance
//     below 0.21 nm (hard-core contacts);
//   * rigid TIP4P/2005 geometry: r_OH 0.09572 nm, HOH 104.52 deg, O-M
//     0.01546 nm on the bisector;
//   * mean-reverting OU motion of the displacement from the lattice site and
//     of the rotation vector from the initial orientation, tau = 100 fs,
//     frame step 10 fs, stationary amplitudes 0.02 nm and 0.3 rad; a
//     molecule's proposed move is rejected (it keeps its previous pose) when
//     it would create a hard-core contact, so no frame holds the unphysical
//     H...H or H...M overlaps whose fields lie far outside the maps' range.
//     Six normals are drawn per molecule and frame whether or not the move is
//     accepted, so the random stream does not depend on the outcome.
// Random numbers: SplitMix64 with Box-Muller normals.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "calcir.h"
#include "common.hpp"

namespace {

struct SplitMix64 {
    uint64_t s;
    bool have_spare = false;
    double spare = 0.0;
    explicit SplitMix64(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return double(next() >> 11) * (1.0 / 9007199254740992.0); }
    double normal() {
        if (have_spare) {
            have_spare = false;
            return spare;
        }
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        const double r = std::sqrt(-2.0 * std::log(u1));
        const double a = 6.283185307179586 * u2;
        spare = r * std::sin(a);
        have_spare = true;
        return r * std::cos(a);
    }
};

struct Mat3 {
    double m[3][3];
};

Mat3 quat_to_mat(double w, double x, double y, double z) {
    Mat3 R;
    R.m[0][0] = 1 - 2 * (y * y + z * z);
    R.m[0][1] = 2 * (x * y - w * z);
    R.m[0][2] = 2 * (x * z + w * y);
    R.m[1][0] = 2 * (x * y + w * z);
    R.m[1][1] = 1 - 2 * (x * x + z * z);
    R.m[1][2] = 2 * (y * z - w * x);
    R.m[2][0] = 2 * (x * z - w * y);
    R.m[2][1] = 2 * (y * z + w * x);
    R.m[2][2] = 1 - 2 * (x * x + y * y);
    return R;
}

// Rodrigues: rotation by the rotation vector phi.
Mat3 rotvec_to_mat(const double* phi) {
    const double th = std::sqrt(phi[0] * phi[0] + phi[1] * phi[1] + phi[2] * phi[2]);
    if (th < 1e-12) return quat_to_mat(1, 0, 0, 0);
    const double h = 0.5 * th, s = std::sin(h) / th;
    return quat_to_mat(std::cos(h), phi[0] * s, phi[1] * s, phi[2] * s);
}

Mat3 mul(const Mat3& A, const Mat3& B) {
    Mat3 C;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double acc = 0;
            for (int k = 0; k < 3; ++k) acc += A.m[i][k] * B.m[k][j];
            C.m[i][j] = acc;
        }
    return C;
}

void apply(const Mat3& R, const double* v, double* out) {
    for (int i = 0; i < 3; ++i) out[i] = R.m[i][0] * v[0] + R.m[i][1] * v[1] + R.m[i][2] * v[2];
}

double mi(double d, double L) { return d - L * std::round(d / L); }

// Body-frame charged sites: H1, H2, M (O at the origin, bisector along +z).
void body_sites(double b[3][3]) {
    const double roh = 0.09572, half = 0.5 * 104.52 * 3.141592653589793 / 180.0, dom = 0.01546;
    b[0][0] = roh * std::sin(half);
    b[0][1] = 0;
    b[0][2] = roh * std::cos(half);
    b[1][0] = -roh * std::sin(half);
    b[1][1] = 0;
    b[1][2] = roh * std::cos(half);
    b[2][0] = 0;
    b[2][1] = 0;
    b[2][2] = dom;
}

}  // namespace

extern "C" {

float cir_synth_box_nm(int nwater) {
    // 128 waters in L = 1.564 nm (0.997 g/cm^3), scaled at fixed density.
    if (nwater <= 0) return 0.0f;
    return float(1.564 * std::cbrt(double(nwater) / 128.0));
}

int cir_synth_trajectory(int nwater, int nframes, uint64_t seed, float box_nm, float* xyz,
                         float* times) try {
    if (nwater <= 0 || nframes <= 0 || !(box_nm > 0.0f) || !xyz) return CIR_ERR_ARG;
    SplitMix64 rng(seed);
    const double L = box_nm;
    int nside = 1;
    while (int64_t(nside) * nside * nside < nwater) ++nside;
    const double a = L / nside;
    const int nsites = nside * nside * nside;
    std::vector<int> order(nsites);
    for (int i = 0; i < nsites; ++i) order[i] = i;
    for (int i = nsites - 1; i > 0; --i) {  // Fisher-Yates
        const int j = int(rng.next() % uint64_t(i + 1));
        std::swap(order[i], order[j]);
    }
    std::vector<double> site(3 * size_t(nwater));
    for (int m = 0; m < nwater; ++m) {
        const int id = order[m];
        const int ix = id % nside, iy = (id / nside) % nside, iz = id / (nside * nside);
        site[3 * m + 0] = (ix + 0.5) * a + (2 * rng.uniform() - 1) * 0.02;
        site[3 * m + 1] = (iy + 0.5) * a + (2 * rng.uniform() - 1) * 0.02;
        site[3 * m + 2] = (iz + 0.5) * a + (2 * rng.uniform() - 1) * 0.02;
    }
    double body[3][3];
    body_sites(body);
    // neighbour lists from the lattice sites (displacements stay << 0.1 nm)
    std::vector<std::vector<int>> nbr(nwater);
    const double rn2 = 0.62 * 0.62;
    for (int m = 0; m < nwater; ++m)
        for (int o = m + 1; o < nwater; ++o) {
            double d2 = 0;
            for (int c = 0; c < 3; ++c) {
                const double d = mi(site[3 * m + c] - site[3 * o + c], L);
                d2 += d * d;
            }
            if (d2 < rn2) {
                nbr[m].push_back(o);
                nbr[o].push_back(m);
            }
        }
    // lab positions of H1, H2, M per molecule
    std::vector<double> pos(9 * size_t(nwater), 0.0);
    auto pose = [&](int m, const Mat3& R, const double* o, double* out9) {
        for (int s2 = 0; s2 < 3; ++s2) {
            double v[3];
            apply(R, body[s2], v);
            for (int c = 0; c < 3; ++c) out9[3 * s2 + c] = o[c] + v[c];
        }
        (void)m;
    };
    const double hh2 = 0.18 * 0.18, hm2 = 0.21 * 0.21;
    auto clashes = [&](int m, const double* p9, const std::vector<char>& placed) {
        for (int o : nbr[m]) {
            if (!placed[o]) continue;
            const double* q9 = &pos[9 * size_t(o)];
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) {
                    if (a == 2 && b == 2) continue;  // M...M carries no H
                    double d2 = 0;
                    for (int c = 0; c < 3; ++c) {
                        const double d = mi(p9[3 * a + c] - q9[3 * b + c], L);
                        d2 += d * d;
                    }
                    const double lim = (a == 2 || b == 2) ? hm2 : hh2;
                    if (d2 < lim) return true;
                }
        }
        return false;
    };
    std::vector<Mat3> R0(nwater);
    std::vector<char> placed(nwater, 0);
    const double tp = 6.283185307179586;
    for (int m = 0; m < nwater; ++m) {
        double p9[9];
        for (int attempt = 0; attempt < 400; ++attempt) {
            const double u1 = rng.uniform(), u2 = rng.uniform(), u3 = rng.uniform();
            const double qx = std::sqrt(1 - u1) * std::sin(tp * u2), qy = std::sqrt(1 - u1) * std::cos(tp * u2);
            const double qz = std::sqrt(u1) * std::sin(tp * u3), qw = std::sqrt(u1) * std::cos(tp * u3);
            R0[m] = quat_to_mat(qw, qx, qy, qz);
            pose(m, R0[m], &site[3 * m], p9);
            if (!clashes(m, p9, placed)) break;
        }
        for (int q = 0; q < 9; ++q) pos[9 * size_t(m) + q] = p9[q];
        placed[m] = 1;
    }
    // OU motion with hard-core rejection
    const double tau = 0.100, dt = 0.010, sx = 0.02, sphi = 0.3;
    const double dec = std::exp(-dt / tau), kick = std::sqrt(1.0 - dec * dec);
    std::vector<double> disp(3 * size_t(nwater), 0.0), phi(3 * size_t(nwater), 0.0);
    std::vector<double> opos(3 * size_t(nwater));  // oxygen positions (unwrapped)
    for (int m = 0; m < nwater; ++m)
        for (int c = 0; c < 3; ++c) opos[3 * m + c] = site[3 * m + c];
    const int natoms = 4 * nwater;
    for (int f = 0; f < nframes; ++f) {
        if (f > 0) {
            for (int m = 0; m < nwater; ++m) {
                double nd[3], np[3], noxy[3], p9[9];
                for (int c = 0; c < 3; ++c) nd[c] = disp[3 * m + c] * dec + sx * kick * rng.normal();
                for (int c = 0; c < 3; ++c) np[c] = phi[3 * m + c] * dec + sphi * kick * rng.normal();
                for (int c = 0; c < 3; ++c) noxy[c] = site[3 * m + c] + nd[c];
                const Mat3 R = mul(rotvec_to_mat(np), R0[m]);
                pose(m, R, noxy, p9);
                if (clashes(m, p9, placed)) continue;  // rejected: keep the previous pose
                for (int c = 0; c < 3; ++c) {
                    disp[3 * m + c] = nd[c];
                    phi[3 * m + c] = np[c];
                    opos[3 * m + c] = noxy[c];
                }
                for (int q = 0; q < 9; ++q) pos[9 * size_t(m) + q] = p9[q];
            }
        }
        float* fr = xyz + size_t(f) * natoms * 3;
        for (int m = 0; m < nwater; ++m) {
            double shift[3];
            for (int c = 0; c < 3; ++c) {
                const double o = opos[3 * m + c];
                shift[c] = -L * std::floor(o / L);  // wrap the oxygen into [0, L), keep the molecule whole
                fr[3 * (4 * m) + c] = float(o + shift[c]);
            }
            for (int s2 = 0; s2 < 3; ++s2)
                for (int c = 0; c < 3; ++c) fr[3 * (4 * m + 1 + s2) + c] = float(pos[9 * size_t(m) + 3 * s2 + c] + shift[c]);
        }
        if (times) times[f] = float(f * 0.010);
    }
    return CIR_OK;
} catch (...) {
    return cir::host_exception("cir_synth_trajectory");
}

}  // extern "C"
