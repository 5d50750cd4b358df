// synth.cpp -- seeded synthetic rigid-water trajectories.
//
// The reference reads GROMACS XTC trajectories (calcIR.cu:212-227, 341-398).
// This source produces frames of the same layout (natoms x rvec, atom order
// O, H1, H2, M per molecule, nm, cubic box, t = k*0.010 ps from 0) from a
// seeded Ornstein-Uhlenbeck rigid-body model, deterministic across machines:
//   * oxygen sites on the smallest simple-cubic lattice with >= n sites,
//     occupied by a seeded shuffle, uniform jitter of +-0.02 nm;
//   * uniform random orientations (Shoemake quaternions), redrawn while any
//     intermolecular H...(H, M) distance is below 0.15 nm;
//   * rigid TIP4P/2005 geometry: r_OH 0.09572 nm, HOH 104.52 deg, O-M
//     0.01546 nm on the bisector;
//   * mean-reverting OU motion of the displacement from the lattice site and
//     of the rotation vector from the initial orientation, tau = 100 fs,
//     frame step 10 fs, stationary amplitudes 0.02 nm and 0.3 rad.
// Random numbers: SplitMix64 with Box-Muller normals.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "calcir.h"

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
                         float* times) {
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
    std::vector<Mat3> R0(nwater);
    std::vector<double> placed(9 * size_t(nwater));  // H1, H2, M lab positions
    const double clash2 = 0.15 * 0.15;
    for (int m = 0; m < nwater; ++m) {
        for (int attempt = 0; attempt < 200; ++attempt) {
            const double u1 = rng.uniform(), u2 = rng.uniform(), u3 = rng.uniform();
            const double tp = 6.283185307179586;
            const double qx = std::sqrt(1 - u1) * std::sin(tp * u2), qy = std::sqrt(1 - u1) * std::cos(tp * u2);
            const double qz = std::sqrt(u1) * std::sin(tp * u3), qw = std::sqrt(u1) * std::cos(tp * u3);
            R0[m] = quat_to_mat(qw, qx, qy, qz);
            for (int s = 0; s < 3; ++s) {
                double v[3];
                apply(R0[m], body[s], v);
                for (int c = 0; c < 3; ++c) placed[9 * m + 3 * s + c] = site[3 * m + c] + v[c];
            }
            bool ok = true;
            for (int o = 0; o < m && ok; ++o) {
                double doo = 0;
                for (int c = 0; c < 3; ++c) {
                    const double d = mi(site[3 * m + c] - site[3 * o + c], L);
                    doo += d * d;
                }
                if (doo > 0.5 * 0.5) continue;
                for (int s1 = 0; s1 < 3 && ok; ++s1)
                    for (int s2 = 0; s2 < 3 && ok; ++s2) {
                        if (s1 == 2 && s2 == 2) continue;  // M...M is not an H contact
                        double dd = 0;
                        for (int c = 0; c < 3; ++c) {
                            const double d = mi(placed[9 * m + 3 * s1 + c] - placed[9 * o + 3 * s2 + c], L);
                            dd += d * d;
                        }
                        if (dd < clash2) ok = false;
                    }
            }
            if (ok) break;
        }
    }
    // OU motion
    const double tau = 0.100, dt = 0.010, sx = 0.02, sphi = 0.3;
    const double dec = std::exp(-dt / tau), kick = std::sqrt(1.0 - dec * dec);
    std::vector<double> disp(3 * size_t(nwater), 0.0), phi(3 * size_t(nwater), 0.0);
    const int natoms = 4 * nwater;
    for (int f = 0; f < nframes; ++f) {
        if (f > 0) {
            for (size_t i = 0; i < disp.size(); ++i) disp[i] = disp[i] * dec + sx * kick * rng.normal();
            for (size_t i = 0; i < phi.size(); ++i) phi[i] = phi[i] * dec + sphi * kick * rng.normal();
        }
        float* fr = xyz + size_t(f) * natoms * 3;
        for (int m = 0; m < nwater; ++m) {
            const Mat3 R = mul(rotvec_to_mat(&phi[3 * m]), R0[m]);
            double o[3];
            for (int c = 0; c < 3; ++c) {
                double p = site[3 * m + c] + disp[3 * m + c];
                p -= L * std::floor(p / L);
                o[c] = p;
                fr[3 * (4 * m) + c] = float(p);
            }
            for (int s = 0; s < 3; ++s) {
                double v[3];
                apply(R, body[s], v);
                for (int c = 0; c < 3; ++c) fr[3 * (4 * m + 1 + s) + c] = float(o[c] + v[c]);
            }
        }
        if (times) times[f] = float(f * 0.010);
    }
    return CIR_OK;
}

}  // extern "C"
