// pipeline.hip -- the calcIR sample x frame pipeline on one GPU
// (calcIR.cu:360-919), as a context bound to one HIP stream.
//
// Two control flows produce the same outputs:
//   CIR_MODE_BATCHED  every distinct trajectory frame needed by any
//     (sample, tcf point) is processed once: field, sites, Hamiltonian and
//     eigensolve run batched over chunks of frames; the IPR histogram and the
//     spectral density are weighted by how many (sample, point) pairs /
//     samples use the frame; all samples active at a frame are propagated in
//     one product in the algebraically identical vector form Y = F M(0)
//     (9 complex columns per sample).
//   CIR_MODE_PLAIN    the reference control flow: sample by sample, frame by
//     frame, one eigensolve per (sample, point) and the literal F-matrix
//     propagation F <- V P V^T F (calcIR.cu:597-617) with the 15 bilinear
//     forms of calcIR.cu:628-745.

#include <algorithm>
#include <atomic>
#include <mutex>
#include <thread>
#include <cstdint>
#include <cstdlib>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"
#include "tuner.hpp"
#include "ctx.hpp"
#include "gemm.hpp"
#include "spectra.hpp"

using namespace cir;

namespace cir {


int build_schedule(const cir_params& p, int nframes, const float* times, Schedule& S) {
    const int ns = p.nsamples, nt = p.ntcfpoints;
    if (nframes < 2) return fail(CIR_ERR_FRAME, "at least two frames are needed to find the frame spacing");
    // calcIR.cu:342-345: the spacing of the first two frames rounded at the
    // trajectory's precision (float product, rounded, divided in double)
    const float prec = p.time_prec;
    if (!(prec > 0.0f)) return fail(CIR_ERR_ARG, "time_precision must be positive");
    const float frame_dt = float(std::round((times[1] - times[0]) * prec) / (1. * prec));
    if (!(frame_dt > 0.0f))
        return fail(CIR_ERR_FRAME, "trajectory times are not increasing at the time precision");
    S.fidx.assign(size_t(ns) * nt, 0);
    for (int s = 0; s < ns; ++s)
        for (int k = 0; k < nt; ++k) {
            float t;
            const int f = cir_frame_index(&p, s, k, frame_dt, &t);
            if (f < 0 || f >= nframes)
                return fail(CIR_ERR_FRAME, "could not find the desired frame at time " + std::to_string(t) + " ps");
            if (std::fabs(t - times[f]) > frame_dt * 1E-1)  // calcIR.cu:391
                return fail(CIR_ERR_FRAME, "could not find the desired frame at time " + std::to_string(t) +
                                               " ps (found " + std::to_string(times[f]) + ")");
            if (k > 0 && f <= S.fidx[size_t(s) * nt + k - 1])
                return fail(CIR_ERR_FRAME, "dt must select increasing trajectory frames");
            S.fidx[size_t(s) * nt + k] = f;
        }
    std::vector<int>& uof = S.uof;
    uof.assign(nframes, -1);
    for (int f : S.fidx) uof[f] = 0;
    S.uniq.clear();
    for (int f = 0; f < nframes; ++f)
        if (uof[f] == 0) {
            uof[f] = int(S.uniq.size());
            S.uniq.push_back(f);
        }
    const int nu = int(S.uniq.size());
    S.mult.assign(nu, 0.0);
    S.orig.assign(nu, 0.0);
    for (int s = 0; s < ns; ++s) {
        for (int k = 0; k < nt; ++k) S.mult[uof[S.fidx[size_t(s) * nt + k]]] += 1.0;
        S.orig[uof[S.fidx[size_t(s) * nt]]] += 1.0;
    }
    // slots: samples in order of their first frame; a slot is held from the
    // first to the last frame of its sample
    std::vector<int> order(ns);
    for (int s = 0; s < ns; ++s) order[s] = s;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return S.fidx[size_t(a) * nt] < S.fidx[size_t(b) * nt];
    });
    std::vector<int> slot_of(ns, -1), slot_end;  // slot_end: last distinct index in use
    for (int s : order) {
        const int u0 = uof[S.fidx[size_t(s) * nt]];
        const int u1 = uof[S.fidx[size_t(s) * nt + nt - 1]];
        int q = -1;
        for (int c = 0; c < int(slot_end.size()); ++c)
            if (slot_end[c] < u0) {
                q = c;
                break;
            }
        if (q < 0) {
            q = int(slot_end.size());
            slot_end.push_back(-1);
        }
        slot_end[q] = u1;
        slot_of[s] = q;
    }
    S.nslot = std::max<int>(1, int(slot_end.size()));
    S.action.assign(size_t(nu) * S.nslot, 0);
    S.target.assign(size_t(nu) * S.nslot, 0);
    for (int s = 0; s < ns; ++s) {
        const int q = slot_of[s];
        const int u0 = uof[S.fidx[size_t(s) * nt]];
        const int u1 = uof[S.fidx[size_t(s) * nt + nt - 1]];
        for (int u = u0; u <= u1; ++u) S.action[size_t(u) * S.nslot + q] = 3;
        for (int k = 0; k < nt; ++k) {
            const int u = uof[S.fidx[size_t(s) * nt + k]];
            S.action[size_t(u) * S.nslot + q] = (k == 0) ? 1 : 2;
            S.target[size_t(u) * S.nslot + q] = s * nt + k;
        }
    }
    return CIR_OK;
}

int finalize(cir_ctx* c, const std::vector<double>& out /* ns*nt*6 */, const std::vector<double>& acc /* 4*nomega */,
             long long skipped) {
    const cir_params& p = c->p;
    const int ns = p.nsamples, nt = p.ntcfpoints;
    // calcIR.cu:646-648, 739-745 (accumulated over samples), 777-781 (decay)
    std::vector<double> X(3 * 2 * size_t(nt), 0.0);
    for (int sp = 0; sp < 3; ++sp) {
        c->tcf[sp].assign(2 * size_t(nt), 0.0);
        for (int k = 0; k < nt; ++k) {
            double re = 0.0, im = 0.0;
            for (int s = 0; s < ns; ++s) {
                re += out[(size_t(s) * nt + k) * 6 + 2 * sp];
                im += out[(size_t(s) * nt + k) * 6 + 2 * sp + 1];
            }
            const double dcy = std::exp(-1.0 * k * p.dt / (2.0 * p.t1)) / (1. * ns);
            c->tcf[sp][2 * k] = re * dcy;
            c->tcf[sp][2 * k + 1] = im * dcy;
            X[(size_t(sp) * nt + k) * 2] = re * dcy;
            X[(size_t(sp) * nt + k) * 2 + 1] = im * dcy;
        }
    }
    // calcIR.cu:788-827: C2R of the zero-padded TCFs
    DevBuf dX, dy;
    CIR_TRY(dX.alloc(X.size() * sizeof(double)));
    CIR_TRY(dy.alloc(3 * size_t(c->nfft) * sizeof(double)));
    CIR_HIP(hipMemcpyAsync(dX.p, X.data(), X.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    {
        ProfScope ps(&c->prof, CIR_PH_FINALIZE, c->stream);
        CIR_TRY(spectrum_c2r(3, c->nfft, nt, dX.as<double>(), dy.as<double>(), c->stream));
    }
    std::vector<double> y(3 * size_t(c->nfft));
    CIR_HIP(hipMemcpyAsync(y.data(), dy.p, y.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    CIR_HIP(hipStreamSynchronize(c->stream));
    for (int sp = 0; sp < 3; ++sp) c->ftcf[sp].assign(y.begin() + size_t(sp) * c->nfft, y.begin() + size_t(sp + 1) * c->nfft);
    // calcIR.cu:831-839
    const int nom = c->nomega;
    const double norm = double(c->N) * ns * nt;
    c->sw.assign(nom, 0.0);
    c->pw.assign(nom, 0.0);
    c->rw.assign(nom, 0.0);
    c->rmw.assign(nom, 0.0);
    for (int i = 0; i < nom; ++i) {
        c->sw[i] = acc[i] / double(ns);
        const double pwi = acc[nom + i] / norm;
        c->pw[i] = pwi;
        c->rw[i] = (acc[2 * nom + i] / norm) / pwi;
        c->rmw[i] = (acc[3 * nom + i] / norm) / pwi;
    }
    c->hist_skipped = skipped;
    c->nsteps = int64_t(ns) * nt;
    c->have = true;
    return CIR_OK;
}

// Gathers trajectory frames frames[0..nb) into contiguous device buffers.
int gather_frames(cir_ctx* c, const float* xyz_d, const float* box_d, const int* frames, int nb, float* xyzc, float* boxc) {
    const size_t fb = size_t(c->natoms) * 3 * sizeof(float);
    const int bs = box_floats(c);
    for (int q = 0; q < nb; ++q) {
        CIR_HIP(hipMemcpyAsync(reinterpret_cast<char*>(xyzc) + q * fb,
                               reinterpret_cast<const char*>(xyz_d) + size_t(frames[q]) * fb, fb, hipMemcpyDeviceToDevice,
                               c->stream));
        CIR_HIP(hipMemcpyAsync(boxc + size_t(bs) * q, box_d + size_t(bs) * frames[q], bs * sizeof(float),
                               hipMemcpyDeviceToDevice, c->stream));
    }
    return CIR_OK;
}

// Host check of the boxes of a trajectory: positive finite lengths; for
// triclinic boxes the GROMACS form (a along x, b in the xy plane) reduced so
// that |b_x|, |c_x| <= a_x/2 and |c_y| <= b_y/2, which the minimum image needs.
int check_boxes(const cir_ctx* c, int nframes, const float* box) {
    const int bs = box_floats(c);
    for (int f = 0; f < nframes; ++f) {
        const float* q = box + size_t(bs) * f;
        for (int i = 0; i < bs; ++i)
            if (!std::isfinite(q[i])) return fail(CIR_ERR_NONFINITE, "non-finite box");
        if (bs == 3) {
            if (!(q[0] > 0) || !(q[1] > 0) || !(q[2] > 0)) return fail(CIR_ERR_NONFINITE, "non-positive box length");
            continue;
        }
        if (!(q[0] > 0) || !(q[4] > 0) || !(q[8] > 0)) return fail(CIR_ERR_NONFINITE, "non-positive box length");
        if (q[1] != 0.0f || q[2] != 0.0f || q[5] != 0.0f)
            return fail(CIR_ERR_ARG, "triclinic box must have a along x and b in the xy plane (frame " +
                                         std::to_string(f) + ")");
        const float tol = 1.0f + 1e-5f;
        if (std::fabs(q[3]) > 0.5f * q[0] * tol || std::fabs(q[6]) > 0.5f * q[0] * tol ||
            std::fabs(q[7]) > 0.5f * q[4] * tol)
            return fail(CIR_ERR_ARG, "triclinic box is not reduced (frame " + std::to_string(f) + ")");
    }
    return CIR_OK;
}

int auto_chunk(const cir_ctx* c, int nuniq) {
    if (c->p.chunk > 0) return std::min(c->p.chunk, nuniq);
    // H and V of every frame of a chunk plus the eigensolver's work space for
    // the chunk, within half of the free device memory (large boxes need
    // large batches: one workgroup per matrix fills the GPU only with
    // hundreds of matrices).  The work space is counted at the chunk's size:
    // its inverse-iteration scratch is capped, not per matrix (at N = 8192,
    // 1.7 GB per matrix, which counted per matrix held the 8192-OH box to
    // 60 matrices per chunk).
    const double hv = 2.0 * double(c->N) * c->ld * c->esz;
    auto need = [&](int B) { return double(B) * hv + double(eigh_workspace(c->real_prec, B, c->N, c->ld)); };
    size_t fr = 0, tot = 0;
    double budget = 24.0 * (1 << 30);
    if (hipMemGetInfo(&fr, &tot) == hipSuccess) budget = std::max(budget, 0.5 * double(fr));
    budget /= std::max(1, c->mem_share);
    int B = std::max(1, std::min(512, nuniq));
    while (B > 1 && need(B) > budget) B = std::max(1, std::min(B - 1, int(double(B) * budget / need(B))));
    // chunks of nearly equal size: a short last chunk would leave most CUs
    // idle in the one-workgroup-per-matrix kernels (and at 16 or fewer large
    // matrices take the one-matrix-at-a-time path)
    const int nchunk = (nuniq + B - 1) / B;
    return (nuniq + nchunk - 1) / nchunk;
}

// ---- checkpoint / resume (batched mode) ------------------------------------
// File <cptf>_ckpt.bin: a header (magic, job key, index of the next distinct
// frame, section sizes), then the propagation state Y (N x 18 nslot reals),
// the TCF accumulators (nsamples x ntcfpoints x 6 doubles), the IPR and
// spectral-density accumulators (4 nomega doubles) and the out-of-window
// count.  The job key hashes the parameters, sizes and frame times; a file
// written for another job is refused.  Writes go to a temporary name and are
// renamed into place.
struct CkptHeader {
    char magic[8];
    uint64_t key;
    int64_t next;
    int64_t ybytes, outbytes, accbytes, g0bytes;
};
constexpr char kCkptMagic[8] = {'C', 'I', 'R', 'C', 'K', 'P', 'T', '2'};

struct Fnv {
    uint64_t h = 1469598103934665603ull;
    void add(const void* p, size_t n) {
        const unsigned char* b = static_cast<const unsigned char*>(p);
        for (size_t q = 0; q < n; ++q) h = (h ^ b[q]) * 1099511628211ull;
    }
    template <typename V>
    void val(const V& v) {
        add(&v, sizeof(v));
    }
};

// The key covers the parameters, the sizes, every frame time and the
// coordinates of the first and last distinct frames (read back from the
// device; hashing every frame would cost a full trajectory copy).
int job_key(const cir_ctx* c, const Schedule& S, int nframes, const float* xyz_d, const float* times,
            uint64_t* out) {
    const cir_params& p = c->p;
    Fnv f;
    f.val(c->N);
    f.val(c->natoms);
    f.val(c->esz);
    f.val(c->imodel);
    f.val(c->ispecies);
    f.val(p.precision);
    f.val(p.ntcfpoints);
    f.val(p.nsamples);
    f.val(p.sampleEvery);
    f.val(p.omegaStep);
    f.val(p.natom_mol);
    f.val(p.nchrom_mol);
    f.val(p.nzeros);
    f.val(p.imap);
    f.val(p.dt);
    f.val(p.t1);
    f.val(p.avef);
    f.val(p.beginTime);
    f.val(p.omegaStart);
    f.val(p.omegaStop);
    f.val(p.cutoff);
    f.val(p.triclinic);
    f.val(p.reunit_every);
    f.val(p.origin_stride);
    f.val(p.oh_per_molecule);
    f.val(p.msite_bisector);
    f.val(p.time_prec);
    f.val(p.deterministic);
    f.val(c->mc);
    f.val(S.nslot);
    f.val(int64_t(S.uniq.size()));
    f.val(nframes);
    f.add(times, sizeof(float) * size_t(nframes));
    const size_t fb = size_t(c->natoms) * 3 * sizeof(float);
    std::vector<unsigned char> fr(fb);
    for (int f0 : {S.uniq.front(), S.uniq.back()}) {
        CIR_HIP(hipMemcpyAsync(fr.data(), reinterpret_cast<const char*>(xyz_d) + size_t(f0) * fb, fb,
                               hipMemcpyDeviceToHost, c->stream));
        CIR_HIP(hipStreamSynchronize(c->stream));
        f.add(fr.data(), fb);
    }
    *out = f.h;
    return CIR_OK;
}

std::string ckpt_path(const cir_params& p) { return std::string(p.cptf) + "_ckpt.bin"; }

int ckpt_write(const std::string& path, uint64_t key, int64_t next, const std::vector<unsigned char>& y,
               const std::vector<double>& out, const std::vector<double>& acc, long long skip,
               const std::vector<double>& g0) {
    CkptHeader h{};
    std::memcpy(h.magic, kCkptMagic, 8);
    h.key = key;
    h.next = next;
    h.ybytes = int64_t(y.size());
    h.outbytes = int64_t(out.size() * sizeof(double));
    h.accbytes = int64_t(acc.size() * sizeof(double));
    h.g0bytes = int64_t(g0.size() * sizeof(double));
    const std::string tmp = path + ".tmp";
    FILE* f = std::fopen(tmp.c_str(), "wb");
    if (!f) return fail(CIR_ERR_IO, "could not open " + tmp + " for the checkpoint");
    bool ok = std::fwrite(&h, sizeof(h), 1, f) == 1;
    ok = ok && std::fwrite(y.data(), 1, y.size(), f) == y.size();
    ok = ok && std::fwrite(out.data(), sizeof(double), out.size(), f) == out.size();
    ok = ok && std::fwrite(acc.data(), sizeof(double), acc.size(), f) == acc.size();
    ok = ok && std::fwrite(&skip, sizeof(skip), 1, f) == 1;
    ok = ok && std::fwrite(g0.data(), sizeof(double), g0.size(), f) == g0.size();
    ok = (std::fclose(f) == 0) && ok;
    if (!ok || std::rename(tmp.c_str(), path.c_str()) != 0) return fail(CIR_ERR_IO, "checkpoint write to " + path + " failed");
    return CIR_OK;
}

int ckpt_read(const std::string& path, uint64_t key, int64_t* next, std::vector<unsigned char>& y,
              std::vector<double>& out, std::vector<double>& acc, long long* skip, std::vector<double>& g0) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return fail(CIR_ERR_IO, "resume: no checkpoint " + path);
    CkptHeader h{};
    bool ok = std::fread(&h, sizeof(h), 1, f) == 1 && std::memcmp(h.magic, kCkptMagic, 8) == 0;
    if (ok && h.key != key) {
        std::fclose(f);
        return fail(CIR_ERR_ARG, "resume: " + path + " was written for another job (parameters or frames differ)");
    }
    ok = ok && h.ybytes == int64_t(y.size()) && h.outbytes == int64_t(out.size() * sizeof(double)) &&
         h.accbytes == int64_t(acc.size() * sizeof(double)) && h.g0bytes == int64_t(g0.size() * sizeof(double));
    ok = ok && std::fread(y.data(), 1, y.size(), f) == y.size();
    ok = ok && std::fread(out.data(), sizeof(double), out.size(), f) == out.size();
    ok = ok && std::fread(acc.data(), sizeof(double), acc.size(), f) == acc.size();
    ok = ok && std::fread(skip, sizeof(*skip), 1, f) == 1;
    ok = ok && std::fread(g0.data(), sizeof(double), g0.size(), f) == g0.size();
    std::fclose(f);
    if (!ok) return fail(CIR_ERR_IO, "resume: " + path + " is not a complete checkpoint");
    *next = h.next;
    return CIR_OK;
}

// A batched run between chunks of distinct frames: the schedule, the device
// buffers, the propagation state Y and the accumulators.  run_batched drives
// it over a trajectory resident in device memory; the streaming entry points
// (cir_stream_*) fill its chunk buffers from host frames as they arrive.
struct BatchRun {
    cir_ctx* c = nullptr;
    Schedule S;
    int nu = 0, B = 0, nslot = 0, ldy = 0, R = 0;
    // One stream: the propagation of a chunk follows its eigensolve.  (Run
    // concurrently on a second stream it only competes with the
    // tridiagonalisation, which keeps every CU busy.)
    DevBuf xyzc, boxc, H, work, wmult, worig, m2, Y0, Y1, G, act, tgt, out, acc, skip, sites, V, w;
    // re-unitarisation: mode per (distinct frame, slot): 1 at an origin
    // (record G0), 2 every reunit_every-th propagated point of a sample
    DevBuf rmode, rg0, rwork;
    DevBuf det;  // fixed-order sums of the histograms and the spectral density (deterministic key)
    std::vector<char> rany;
    void* Ycur = nullptr;
    void* Ynext = nullptr;

    int init(cir_ctx* ctx, int nframes, const float* times) {
        c = ctx;
        const cir_params& p = c->p;
        CIR_TRY(build_schedule(p, nframes, times, S));
        nu = int(S.uniq.size());
        const int N = c->N, ld = c->ld, ns = p.nsamples, nt = p.ntcfpoints, nom = c->nomega;
        B = auto_chunk(c, nu);
        const size_t esz = c->esz;
        hipStream_t st = c->stream;
        CIR_TRY(xyzc.alloc(size_t(B) * c->natoms * 3 * sizeof(float)));
        CIR_TRY(boxc.alloc(size_t(B) * box_floats(c) * sizeof(float)));
        CIR_TRY(sites.alloc(size_t(B) * CIR_NSITE * N * esz));
        CIR_TRY(V.alloc(size_t(B) * N * ld * esz));
        CIR_TRY(w.alloc(size_t(B) * N * sizeof(double)));
        CIR_TRY(H.alloc(size_t(B) * N * ld * esz));
        CIR_TRY(work.alloc(eigh_workspace(c->real_prec, B, N, ld)));
        CIR_TRY(wmult.alloc(size_t(nu) * sizeof(double)));
        CIR_TRY(worig.alloc(size_t(nu) * sizeof(double)));
        CIR_TRY(m2.alloc(size_t(B) * N * sizeof(double)));
        nslot = S.nslot;
        ldy = 18 * nslot;
        CIR_TRY(Y0.alloc(size_t(N) * ldy * esz));
        CIR_TRY(Y1.alloc(size_t(N) * ldy * esz));
        CIR_TRY(G.alloc(propagate_work_elems(N, 9 * nslot) * esz));
        CIR_TRY(act.alloc(S.action.size() * sizeof(int)));
        CIR_TRY(tgt.alloc(S.target.size() * sizeof(int)));
        CIR_TRY(out.alloc(size_t(ns) * nt * 6 * sizeof(double)));
        CIR_TRY(acc.alloc(4 * size_t(nom) * sizeof(double)));
        CIR_TRY(skip.alloc(sizeof(long long)));
        if (p.deterministic)
            CIR_TRY(det.alloc(std::max(ipr_hist_det_bytes(B, N), spectral_density_det_bytes(B, nom))));
        R = p.reunit_every;
        if (R > 0) {
            std::vector<int> hmode(size_t(nu) * nslot, 0);
            rany.assign(nu, 0);
            for (int u = 0; u < nu; ++u)
                for (int q = 0; q < nslot; ++q) {
                    const size_t e = size_t(u) * nslot + q;
                    const int k = S.target[e] % nt;
                    if (S.action[e] == 1) hmode[e] = 1;
                    if (S.action[e] == 2 && k % R == 0) hmode[e] = 2;
                    rany[u] |= hmode[e] != 0;
                }
            CIR_TRY(rmode.alloc(hmode.size() * sizeof(int)));
            CIR_TRY(rg0.alloc(size_t(nslot) * 81 * sizeof(double)));
            CIR_TRY(rwork.alloc(reunit_work_doubles(nslot) * sizeof(double)));
            CIR_HIP(hipMemcpyAsync(rmode.p, hmode.data(), rmode.bytes, hipMemcpyHostToDevice, st));
            CIR_HIP(hipMemsetAsync(rg0.p, 0, rg0.bytes, st));
            // the host copy must outlive the asynchronous copy
            CIR_HIP(hipStreamSynchronize(st));
        }
        CIR_HIP(hipMemcpyAsync(wmult.p, S.mult.data(), nu * sizeof(double), hipMemcpyHostToDevice, st));
        CIR_HIP(hipMemcpyAsync(worig.p, S.orig.data(), nu * sizeof(double), hipMemcpyHostToDevice, st));
        CIR_HIP(hipMemcpyAsync(act.p, S.action.data(), S.action.size() * sizeof(int), hipMemcpyHostToDevice, st));
        CIR_HIP(hipMemcpyAsync(tgt.p, S.target.data(), S.target.size() * sizeof(int), hipMemcpyHostToDevice, st));
        CIR_HIP(hipMemsetAsync(out.p, 0, out.bytes, st));
        CIR_HIP(hipMemsetAsync(acc.p, 0, acc.bytes, st));
        CIR_HIP(hipMemsetAsync(skip.p, 0, skip.bytes, st));
        CIR_HIP(hipMemsetAsync(Y0.p, 0, Y0.bytes, st));
        CIR_HIP(hipMemsetAsync(Y1.p, 0, Y1.bytes, st));
        Ycur = Y0.p;
        Ynext = Y1.p;
        return CIR_OK;
    }

    // Distinct frames c0 .. c0+nb-1, whose coordinates and boxes are in
    // xyzc / boxc: field, Hamiltonian, eigensystems, histograms and spectral
    // density, then frame by frame propagation and correlations.
    int chunk(int c0, int nb) {
        const cir_params& p = c->p;
        const int N = c->N, ld = c->ld, nom = c->nomega;
        const size_t esz = c->esz;
        const int rp = c->real_prec;
        hipStream_t st = c->stream;
        Prof* pf = &c->prof;
        double* accp = acc.as<double>();
        {
            ProfScope ps(pf, CIR_PH_SITES, st);
            CIR_TRY(eproj_sites(rp, nb, c->natoms, p.natom_mol, p.nchrom_mol, c->imodel, site_rules(p), c->mc, p.cutoff,
                                xyzc.as<float>(), boxc.as<float>(), c->boxk, sites.p, st));
        }
        {
            ProfScope ps(pf, CIR_PH_HAMILTONIAN, st);
            CIR_TRY(hamiltonian(rp, nb, N, ld, p.nchrom_mol, c->ispecies, c->mc, p.avef, boxc.as<float>(), c->boxk,
                                sites.p, H.p, st));
        }
        CIR_TRY(eigh_batched(c->prop_prec, nb, N, ld, H.p, V.p, w.as<double>(), work.p, st, pf));
        {
            ProfScope ps(pf, CIR_PH_IPR, st);
            CIR_TRY(ipr_hist(rp, nb, N, ld, p.nchrom_mol, p.avef, p.omegaStart, p.omegaStep, nom, V.p, w.as<double>(),
                             wmult.as<double>() + c0, accp + nom, accp + 2 * nom, accp + 3 * nom,
                             skip.as<long long>(), st, det.p));
        }
        {
            ProfScope ps(pf, CIR_PH_SW, st);
            CIR_TRY(spectral_density(rp, nb, N, ld, p.t1, p.avef, p.omegaStart, p.omegaStep, nom, V.p,
                                     w.as<double>(), sites.p, worig.as<double>() + c0, accp, m2.as<double>(), st,
                                     det.p));
        }
        // frame by frame: propagation, origin bookkeeping, re-unitarisation,
        // correlations
        auto frame = [&](int u, bool may_tune, hipStream_t st) -> int {
            const int bl = u - c0;
            const char* Vu = V.as<char>() + size_t(bl) * N * ld * esz;
            const double* wu = w.as<double>() + size_t(bl) * N;
            const char* su = sites.as<char>() + size_t(bl) * CIR_NSITE * N * esz;
            bool has_prop = false;
            for (int q = 0; q < nslot; ++q) has_prop |= S.action[size_t(u) * nslot + q] == 2;
            if (has_prop) {
                ProfScope ps(pf, CIR_PH_PROPAGATE, st);
                CIR_TRY(propagate(rp, c->prop_prec, N, ld, 9 * nslot, ldy, double(p.dt), Vu, wu, Ycur, Ynext, G.p,
                                  st, may_tune));
            }
            {
                ProfScope ps(pf, CIR_PH_CORRELATE, st);
                const int* au = act.as<int>() + size_t(u) * nslot;
                const int* tu = tgt.as<int>() + size_t(u) * nslot;
                if (R > 0 && rany[u]) {
                    CIR_TRY(slot_fixup(rp, N, nslot, ldy, au, su, Ycur, Ynext, st));
                    CIR_TRY(reunitarise(rp, N, nslot, ldy, rmode.as<int>() + size_t(u) * nslot, rg0.as<double>(),
                                        Ynext, rwork.as<double>(), st));
                    CIR_TRY(correlate(rp, N, nslot, ldy, au, tu, su, Ynext, out.as<double>(), st));
                } else {
                    CIR_TRY(slot_correlate(rp, N, nslot, ldy, au, tu, su, Ycur, Ynext, out.as<double>(), st));
                }
            }
            std::swap(Ycur, Ynext);
            return CIR_OK;
        };
        int u = c0;
        const int uend = c0 + nb;
        if (use_graphs()) {
            // the frames run directly until the propagation's split is
            // settled (its first call times candidates), the rest of the chunk
            // is captured into one graph and launched at once
            while (u < uend && propagate_needs_tuning(rp, c->prop_prec, N, 9 * nslot)) CIR_TRY(frame(u++, true, st));
            if (uend - u >= 2) {
                // recorded on a stream of its own (the context's stream may be
                // the null stream, which cannot be captured), launched on the
                // context's stream
                if (!cap) CIR_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
                hipGraph_t g = nullptr;
                CIR_HIP(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
                int rc = CIR_OK;
                for (; u < uend && rc == CIR_OK; ++u) rc = frame(u, false, cap);
                const hipError_t e = hipStreamEndCapture(cap, &g);
                if (rc != CIR_OK) {
                    if (e == hipSuccess && g) (void)hipGraphDestroy(g);
                    return rc;
                }
                CIR_HIP(e);
                hipGraphExec_t ge = nullptr;
                const hipError_t ei = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                CIR_HIP(ei);
                graphs.push_back(ge);  // destroyed once the stream has passed it
                CIR_HIP(hipGraphLaunch(ge, st));
            }
        }
        for (; u < uend; ++u) CIR_TRY(frame(u, true, st));
        return CIR_OK;
    }

    // Whether the frame loops run as graphs: not while phases are timed (the
    // per-phase events need the launches one by one), and not under
    // CIR_NO_GRAPHS=1 in the environment.
    bool use_graphs() const {
        if (c->prof.on) return false;
        const char* e = std::getenv("CIR_NO_GRAPHS");
        return !(e && std::atoi(e) == 1);
    }
    std::vector<hipGraphExec_t> graphs;
    hipStream_t cap = nullptr;  // capture stream (records, never runs work)
    ~BatchRun() {
        if (!graphs.empty() && c) (void)hipStreamSynchronize(c->stream);
        for (hipGraphExec_t ge : graphs) (void)hipGraphExecDestroy(ge);
        if (cap) (void)hipStreamDestroy(cap);
    }

    // Host copies of the state a checkpoint holds.
    int save(std::vector<unsigned char>& hy, std::vector<double>& ho, std::vector<double>& ha, long long& hs,
             std::vector<double>& hg) {
        hipStream_t st = c->stream;
        hy.resize(Y0.bytes);
        ho.resize(out.bytes / sizeof(double));
        ha.resize(acc.bytes / sizeof(double));
        hg.resize(R > 0 ? size_t(nslot) * 81 : 0);
        if (R > 0) CIR_HIP(hipMemcpyAsync(hg.data(), rg0.p, rg0.bytes, hipMemcpyDeviceToHost, st));
        CIR_HIP(hipMemcpyAsync(hy.data(), Ycur, hy.size(), hipMemcpyDeviceToHost, st));
        CIR_HIP(hipMemcpyAsync(ho.data(), out.p, out.bytes, hipMemcpyDeviceToHost, st));
        CIR_HIP(hipMemcpyAsync(ha.data(), acc.p, acc.bytes, hipMemcpyDeviceToHost, st));
        CIR_HIP(hipMemcpyAsync(&hs, skip.p, sizeof(hs), hipMemcpyDeviceToHost, st));
        CIR_HIP(hipStreamSynchronize(st));
        return CIR_OK;
    }

    // The accumulators to the host and the outputs into the context.
    int finish() {
        hipStream_t st = c->stream;
        std::vector<double> hout(out.bytes / sizeof(double)), hacc(acc.bytes / sizeof(double));
        long long hskip = 0;
        CIR_HIP(hipMemcpyAsync(hout.data(), out.p, out.bytes, hipMemcpyDeviceToHost, st));
        CIR_HIP(hipMemcpyAsync(hacc.data(), acc.p, acc.bytes, hipMemcpyDeviceToHost, st));
        CIR_HIP(hipMemcpyAsync(&hskip, skip.p, sizeof(long long), hipMemcpyDeviceToHost, st));
        CIR_HIP(hipStreamSynchronize(st));
        c->nframes_used = nu;
        return finalize(c, hout, hacc, hskip);
    }
};

int run_batched(cir_ctx* c, int nframes, const float* xyz_d, const float* box_d, const float* times) {
    const cir_params& p = c->p;
    BatchRun r;
    CIR_TRY(r.init(c, nframes, times));
    const int nu = r.nu, B = r.B;
    hipStream_t st = c->stream;
    const bool use_ckpt = p.ckpt_every > 0 || p.resume != 0 || p.max_frames > 0;
    uint64_t key = 0;
    if (use_ckpt) CIR_TRY(job_key(c, r.S, nframes, xyz_d, times, &key));
    int start = 0;
    if (p.resume) {
        std::vector<unsigned char> hy(r.Y0.bytes);
        std::vector<double> ho(r.out.bytes / sizeof(double)), ha(r.acc.bytes / sizeof(double));
        std::vector<double> hg(r.R > 0 ? size_t(r.nslot) * 81 : 0);
        long long hs = 0;
        int64_t next = 0;
        CIR_TRY(ckpt_read(ckpt_path(p), key, &next, hy, ho, ha, &hs, hg));
        if (next < 0 || next > nu) return fail(CIR_ERR_IO, "resume: checkpoint position out of range");
        if (r.R > 0) CIR_HIP(hipMemcpyAsync(r.rg0.p, hg.data(), r.rg0.bytes, hipMemcpyHostToDevice, st));
        CIR_HIP(hipMemcpyAsync(r.Y0.p, hy.data(), hy.size(), hipMemcpyHostToDevice, st));
        CIR_HIP(hipMemcpyAsync(r.out.p, ho.data(), r.out.bytes, hipMemcpyHostToDevice, st));
        CIR_HIP(hipMemcpyAsync(r.acc.p, ha.data(), r.acc.bytes, hipMemcpyHostToDevice, st));
        CIR_HIP(hipMemcpyAsync(r.skip.p, &hs, sizeof(hs), hipMemcpyHostToDevice, st));
        CIR_HIP(hipStreamSynchronize(st));
        start = int(next);
    }
    long long done_here = 0, since_ckpt = 0;
    for (int c0 = start; c0 < nu; c0 += B) {
        const int nb = std::min(B, nu - c0);
        CIR_TRY(gather_frames(c, xyz_d, box_d, r.S.uniq.data() + c0, nb, r.xyzc.as<float>(), r.boxc.as<float>()));
        CIR_TRY(r.chunk(c0, nb));
        done_here += nb;
        since_ckpt += nb;
        const bool more = c0 + nb < nu;
        const bool halt = more && p.max_frames > 0 && done_here >= p.max_frames;
        if (more && (halt || (p.ckpt_every > 0 && since_ckpt >= p.ckpt_every))) {
            std::vector<unsigned char> hy;
            std::vector<double> ho, ha, hg;
            long long hs = 0;
            CIR_TRY(r.save(hy, ho, ha, hs, hg));
            CIR_TRY(ckpt_write(ckpt_path(p), key, c0 + nb, hy, ho, ha, hs, hg));
            since_ckpt = 0;
        }
        if (halt) {
            set_error("stopped after " + std::to_string(done_here) + " distinct frames; checkpoint " + ckpt_path(p));
            return CIR_INCOMPLETE;
        }
    }
    return r.finish();
}

// A batched run fed from the host: the schedule comes from the frame times
// given up front (calcIR.cu:360-395 fixes the frames from the times alone),
// the frames then arrive in trajectory order, only those the schedule uses are
// copied into the chunk buffers, and a chunk runs as soon as its last distinct
// frame has arrived.  The trajectory never has to be resident, as the
// reference's on-demand xdr_seek/read_xtc reads (calcIR.cu:380-386) allow.
struct StreamRun {
    BatchRun run;
    int nframes = 0;
    int next = 0;  // trajectory index of the next frame to arrive
    int c0 = 0;    // first distinct frame of the chunk being filled
    hipEvent_t e0 = nullptr;
    ~StreamRun() {
        if (e0) (void)hipEventDestroy(e0);
    }
};
void stream_release(StreamRun* s) { delete s; }

// A captured eigensolve of the plain control flow, released after its stream
// has passed it.
struct PlainGraph {
    hipGraphExec_t exec = nullptr;
    hipStream_t cap = nullptr, stream = nullptr;
    ~PlainGraph() {
        if (exec) {
            (void)hipStreamSynchronize(stream);
            (void)hipGraphExecDestroy(exec);
        }
        if (cap) (void)hipStreamDestroy(cap);
    }
};

int run_plain(cir_ctx* c, int nframes, const float* xyz_d, const float* box_d, const float* times) {
    const cir_params& p = c->p;
    if (p.ckpt_every > 0 || p.resume != 0 || p.max_frames > 0)
        return fail(CIR_ERR_ARG, "checkpoint_every, resume and max_frames need mode batched");
    if (p.reunit_every > 0) return fail(CIR_ERR_ARG, "reunitarise_every needs mode batched");
    Schedule S;
    CIR_TRY(build_schedule(p, nframes, times, S));
    const int N = c->N, ld = c->ld, ns = p.nsamples, nt = p.ntcfpoints, nom = c->nomega;
    const size_t esz = c->esz;
    const int rp = c->real_prec;
    hipStream_t st = c->stream;
    const int ldf = 2 * N;
    DevBuf xyz1, box1, sites1, sites0, H, V, w, work, m2, F, G, Y, out, acc, skip, det;
    CIR_TRY(xyz1.alloc(size_t(c->natoms) * 3 * sizeof(float)));
    CIR_TRY(box1.alloc(box_floats(c) * sizeof(float)));
    CIR_TRY(sites1.alloc(size_t(CIR_NSITE) * N * esz));
    CIR_TRY(sites0.alloc(size_t(CIR_NSITE) * N * esz));
    CIR_TRY(H.alloc(size_t(N) * ld * esz));
    CIR_TRY(V.alloc(size_t(N) * ld * esz));
    CIR_TRY(w.alloc(size_t(N) * sizeof(double)));
    CIR_TRY(work.alloc(eigh_workspace(rp, 1, N, ld)));
    CIR_TRY(m2.alloc(size_t(N) * sizeof(double)));
    CIR_TRY(F.alloc(size_t(N) * ldf * esz));
    CIR_TRY(G.alloc(propagate_work_elems(N, N) * esz));
    CIR_TRY(Y.alloc(size_t(N) * 18 * esz));
    CIR_TRY(out.alloc(size_t(ns) * nt * 6 * sizeof(double)));
    CIR_TRY(acc.alloc(4 * size_t(nom) * sizeof(double)));
    CIR_TRY(skip.alloc(sizeof(long long)));
    if (p.deterministic) CIR_TRY(det.alloc(std::max(ipr_hist_det_bytes(1, N), spectral_density_det_bytes(1, nom))));
    CIR_HIP(hipMemsetAsync(out.p, 0, out.bytes, st));
    CIR_HIP(hipMemsetAsync(acc.p, 0, acc.bytes, st));
    CIR_HIP(hipMemsetAsync(skip.p, 0, skip.bytes, st));
    double* accp = acc.as<double>();
    // The eigensolve of every step reads and writes the same buffers with the
    // same launches (one matrix across the GPU: a few launches per
    // reflector), so its second call is captured into a graph and later steps
    // replay it (the first call sets kernel attributes and times its product
    // tiles).  Not while phases are timed, nor under CIR_NO_GRAPHS=1.
    const char* ng = std::getenv("CIR_NO_GRAPHS");
    const bool graphs = !c->prof.on && !(ng && std::atoi(ng) == 1);
    PlainGraph eg;
    int ecalls = 0;
    auto eigh_step = [&]() -> int {
        if (!graphs || ecalls++ == 0) return eigh_batched(c->prop_prec, 1, N, ld, H.p, V.p, w.as<double>(), work.p, st,
                                                         &c->prof);
        if (!eg.exec) {
            CIR_HIP(hipStreamCreateWithFlags(&eg.cap, hipStreamNonBlocking));
            hipGraph_t g = nullptr;
            CIR_HIP(hipStreamBeginCapture(eg.cap, hipStreamCaptureModeThreadLocal));
            const int rc = eigh_batched(c->prop_prec, 1, N, ld, H.p, V.p, w.as<double>(), work.p, eg.cap, nullptr);
            const hipError_t e = hipStreamEndCapture(eg.cap, &g);
            if (rc != CIR_OK) {
                if (e == hipSuccess && g) (void)hipGraphDestroy(g);
                return rc;
            }
            CIR_HIP(e);
            const hipError_t ei = hipGraphInstantiate(&eg.exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            CIR_HIP(ei);
            eg.stream = st;
        }
        CIR_HIP(hipGraphLaunch(eg.exec, st));
        return CIR_OK;
    };
    for (int s = 0; s < ns; ++s) {
        for (int k = 0; k < nt; ++k) {
            const int f = S.fidx[size_t(s) * nt + k];
            CIR_TRY(gather_frames(c, xyz_d, box_d, &f, 1, xyz1.as<float>(), box1.as<float>()));
            Prof* pf = &c->prof;
            {
                ProfScope ps(pf, CIR_PH_SITES, st);
                CIR_TRY(eproj_sites(rp, 1, c->natoms, p.natom_mol, p.nchrom_mol, c->imodel, site_rules(p), c->mc,
                                        p.cutoff, xyz1.as<float>(), box1.as<float>(), c->boxk, sites1.p, st));
            }
            {
                ProfScope ps(pf, CIR_PH_HAMILTONIAN, st);
                CIR_TRY(hamiltonian(rp, 1, N, ld, p.nchrom_mol, c->ispecies, c->mc, p.avef, box1.as<float>(), c->boxk,
                                        sites1.p, H.p, st));
            }
            CIR_TRY(eigh_step());
            {
                ProfScope ps(pf, CIR_PH_IPR, st);
                CIR_TRY(ipr_hist(rp, 1, N, ld, p.nchrom_mol, p.avef, p.omegaStart, p.omegaStep, nom, V.p,
                                 w.as<double>(), nullptr, accp + nom, accp + 2 * nom, accp + 3 * nom,
                                 skip.as<long long>(), st, det.p));
            }
            if (k == 0) {
                ProfScope ps(pf, CIR_PH_SW, st);
                CIR_TRY(spectral_density(rp, 1, N, ld, p.t1, p.avef, p.omegaStart, p.omegaStep, nom, V.p,
                                         w.as<double>(), sites1.p, nullptr, accp, m2.as<double>(), st, det.p));
                CIR_TRY(set_identity(rp, N, ldf, F.p, st));
                CIR_HIP(hipMemcpyAsync(sites0.p, sites1.p, sites1.bytes, hipMemcpyDeviceToDevice, st));
            } else {
                ProfScope ps(pf, CIR_PH_PROPAGATE, st);
                CIR_TRY(propagate(rp, c->prop_prec, N, ld, N, ldf, double(p.dt), V.p, w.as<double>(), F.p, F.p, G.p,
                                  st));
            }
            ProfScope pcorr(pf, CIR_PH_CORRELATE, st);
            // Y = F M(0): Re and Im blocks
            for (int part = 0; part < 2; ++part) {
                GemmArgs g;
                g.M = N;
                g.N = 9;
                g.K = N;
                g.A = F.as<char>() + size_t(part) * N * esz;
                g.a_m = ldf;
                g.a_k = 1;
                g.B = sites0.as<char>() + size_t(CIR_SITE_MU) * N * esz;
                g.b_k = 1;
                g.b_n = N;
                g.C = Y.as<char>() + size_t(part) * 9 * esz;
                g.c_m = 18;
                g.c_n = 1;
                CIR_TRY(gemm(c->prop_prec, g, st));
            }
            CIR_TRY(correlate(rp, N, 1, 18, nullptr, nullptr, sites1.p, Y.p,
                              out.as<double>() + 6 * (size_t(s) * nt + k), st));
        }
    }
    std::vector<double> hout(size_t(ns) * nt * 6), hacc(4 * size_t(nom));
    long long hskip = 0;
    CIR_HIP(hipMemcpyAsync(hout.data(), out.p, out.bytes, hipMemcpyDeviceToHost, st));
    CIR_HIP(hipMemcpyAsync(hacc.data(), acc.p, acc.bytes, hipMemcpyDeviceToHost, st));
    CIR_HIP(hipMemcpyAsync(&hskip, skip.p, sizeof(long long), hipMemcpyDeviceToHost, st));
    CIR_HIP(hipStreamSynchronize(st));
    c->nframes_used = int64_t(ns) * nt;
    return finalize(c, hout, hacc, hskip);
}

}  // namespace

extern "C" int cir_device_count(int* n) {
    if (!n) return fail(CIR_ERR_ARG, "null argument");
    int cnt = 0;
    hipError_t e = hipGetDeviceCount(&cnt);
    if (e != hipSuccess) {
        *n = 0;
        return hip_fail(e, "hipGetDeviceCount", __FILE__, __LINE__);
    }
    *n = cnt;
    return CIR_OK;
}

extern "C" int cir_create(const cir_params* p, int natoms, void* stream, cir_ctx** out) try {
    if (!p || !out) return fail(CIR_ERR_ARG, "cir_create: null argument");
    *out = nullptr;
    int imodel = 0, ispecies = 0;
    CIR_TRY(cir_resolve_model(p->model, &imodel));
    CIR_TRY(cir_resolve_species(p->species, &ispecies));
    MapCoef mc;
    if (!map_coef(p->imap, ispecies, &mc)) return fail(CIR_ERR_MAP, "map must be 0 or 1");
    if (p->natom_mol < 3 || p->nchrom_mol < 1 || natoms <= 0 || natoms % p->natom_mol)
        return fail(CIR_ERR_ARG, "natoms must be a positive multiple of natom_mol (>= 3)");
    if (p->oh_per_molecule != 0 && p->oh_per_molecule != 1) return fail(CIR_ERR_ARG, "oh_per_molecule must be 0 or 1");
    if (p->msite_bisector != 0 && p->msite_bisector != 1) return fail(CIR_ERR_ARG, "msite_bisector must be 0 or 1");
    if (p->oh_per_molecule && p->nchrom_mol > 2)
        return fail(CIR_ERR_ARG, "oh_per_molecule needs nchrom_mol 1 or 2 (OH stretches of a water)");
    if (p->triclinic != 0 && p->triclinic != 1) return fail(CIR_ERR_ARG, "triclinic must be 0 or 1");
    if (p->reunit_every < 0) return fail(CIR_ERR_ARG, "reunitarise_every must be >= 0");
    if (p->origin_stride < 0) return fail(CIR_ERR_ARG, "origin_stride_frames must be >= 0");
    if (!(p->time_prec > 0.0f) || !std::isfinite(p->time_prec))
        return fail(CIR_ERR_ARG, "time_precision must be positive");
    if (p->ntcfpoints <= 0 || p->nsamples <= 0 || p->omegaStep <= 0 || p->nzeros < 0 || !(p->t1 > 0) || !(p->dt > 0))
        return fail(CIR_ERR_ARG, "ntcfpoints, nsamples, omegaStep, t1 and dt must be positive, nzeros >= 0");
    if (cir_nomega(p) <= 0) return fail(CIR_ERR_ARG, "omegaStop must not be below omegaStart");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(CIR_ERR_NODEVICE, "no HIP device available");
    cir_ctx* c = new cir_ctx();
    c->p = *p;
    c->imodel = imodel;
    c->ispecies = ispecies;
    c->mc = mc;
    c->boxk = p->triclinic ? 1 : 0;
    c->natoms = natoms;
    c->nmol = natoms / p->natom_mol;
    c->N = c->nmol * p->nchrom_mol;
    c->ld = (c->N + 3) & ~3;
    c->nomega = cir_nomega(p);
    c->nfft = p->ntcfpoints + p->nzeros;
    c->real_prec = (p->precision == CIR_PREC_FP64) ? CIR_PREC_FP64 : CIR_PREC_FP32;
    c->prop_prec = p->precision;
    c->esz = (c->real_prec == CIR_PREC_FP64) ? sizeof(double) : sizeof(float);
    c->stream = as_stream(stream);
    if (p->deterministic != 0 && p->deterministic != 1) {
        delete c;
        return fail(CIR_ERR_ARG, "deterministic must be 0 or 1");
    }
    if (p->deterministic && (c->real_prec == CIR_PREC_FP64 || !deterministic_eigh_size(c->N))) {
        const int n = c->N;
        delete c;
        return fail(CIR_ERR_ARG, "deterministic needs precision fp32 or bf16x3 and 66 <= N <= " +
                                     std::to_string(deterministic_eigh_max()) + " (N = " + std::to_string(n) + ")");
    }
    *out = c;
    return CIR_OK;
} catch (...) {
    return cir::host_exception("cir_create");
}

namespace {
void reset_phases(cir_ctx* c) {
    for (int i = 0; i < CIR_NPHASE; ++i) {
        c->phase_sec[i] = 0.0;
        c->phase_cnt[i] = 0;
    }
}
}  // namespace

extern "C" int cir_run_device(cir_ctx* c, int nframes, const float* xyz_d, const float* box_d, const float* times) try {
    if (!c || !xyz_d || !box_d || !times || nframes <= 0) return fail(CIR_ERR_ARG, "cir_run_device: bad arguments");
    {
        // the boxes are few bytes: check them on the host
        std::vector<float> hb(size_t(nframes) * box_floats(c));
        CIR_HIP(hipMemcpyAsync(hb.data(), box_d, hb.size() * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        CIR_HIP(hipStreamSynchronize(c->stream));
        CIR_TRY(check_boxes(c, nframes, hb.data()));
    }
    c->have = false;
    reset_phases(c);
    DeterministicScope det(c->p.deterministic != 0);
    EventPair ev;
    CIR_HIP(ev.create());
    CIR_HIP(hipEventRecord(ev.a, c->stream));
    int st = (c->p.mode == CIR_MODE_PLAIN) ? run_plain(c, nframes, xyz_d, box_d, times)
                                           : run_batched(c, nframes, xyz_d, box_d, times);
    if (hipEventRecord(ev.b, c->stream) == hipSuccess && hipEventSynchronize(ev.b) == hipSuccess) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, ev.a, ev.b) == hipSuccess) c->seconds = ms * 1e-3;
    }
    c->prof.collect(c->phase_sec, c->phase_cnt, CIR_NPHASE);
    return st;
} catch (...) {
    return cir::host_exception("cir_run_device");
}

extern "C" int cir_stream_begin(cir_ctx* c, int nframes, const float* times) try {
    if (!c || !times || nframes <= 0) return fail(CIR_ERR_ARG, "cir_stream_begin: bad arguments");
    c->streamed.reset();
    if (c->p.mode == CIR_MODE_PLAIN) return fail(CIR_ERR_ARG, "cir_stream_begin: streaming needs mode batched");
    if (c->p.ckpt_every > 0 || c->p.resume != 0 || c->p.max_frames > 0)
        return fail(CIR_ERR_ARG,
                    "cir_stream_begin: checkpoint_every, resume and max_frames need the whole trajectory (cir_run)");
    c->have = false;
    reset_phases(c);
    DeterministicScope det(c->p.deterministic != 0);
    std::unique_ptr<StreamRun, StreamDeleter> s(new StreamRun());
    CIR_HIP(hipEventCreate(&s->e0));
    CIR_HIP(hipEventRecord(s->e0, c->stream));
    CIR_TRY(s->run.init(c, nframes, times));
    s->nframes = nframes;
    c->streamed = std::move(s);
    return CIR_OK;
} catch (...) {
    return cir::host_exception("cir_stream_begin");
}

namespace {
int stream_push(cir_ctx* c, int first, int count, const float* xyz, const float* box) {
    if (count < 0 || (count > 0 && (!xyz || !box))) return fail(CIR_ERR_ARG, "cir_stream_push: bad arguments");
    StreamRun* s = c->streamed.get();
    if (!s) return fail(CIR_ERR_ARG, "cir_stream_push: no stream is open (cir_stream_begin)");
    if (first != s->next)
        return fail(CIR_ERR_ARG, "cir_stream_push: frames arrive in trajectory order; expected frame " +
                                     std::to_string(s->next) + ", got " + std::to_string(first));
    if (count > s->nframes - first) return fail(CIR_ERR_ARG, "cir_stream_push: frames past the end of the trajectory");
    const size_t fx = size_t(c->natoms) * 3;
    const int bs = box_floats(c);
    for (size_t i = 0; i < fx * count; ++i)
        if (!std::isfinite(xyz[i])) return fail(CIR_ERR_NONFINITE, "non-finite coordinate in the trajectory");
    CIR_TRY(check_boxes(c, count, box));
    DeterministicScope det(c->p.deterministic != 0);
    BatchRun& r = s->run;
    int st = CIR_OK;
    for (int k = 0; k < count && st == CIR_OK; ++k) {
        const int u = r.S.uof[size_t(first) + k];
        if (u < 0) continue;
        // distinct frames arrive in ascending order and a chunk runs when its
        // last one arrives, so this one is at 0 <= slot < B
        const int slot = u - s->c0;
        CIR_HIP(hipMemcpyAsync(r.xyzc.as<float>() + size_t(slot) * fx, xyz + size_t(k) * fx, fx * sizeof(float),
                               hipMemcpyHostToDevice, c->stream));
        CIR_HIP(hipMemcpyAsync(r.boxc.as<float>() + size_t(slot) * bs, box + size_t(k) * bs, bs * sizeof(float),
                               hipMemcpyHostToDevice, c->stream));
        const int nb = std::min(r.B, r.nu - s->c0);
        if (slot == nb - 1) {
            st = r.chunk(s->c0, nb);
            s->c0 += nb;
        }
    }
    s->next = first + count;
    // the caller may reuse its arrays once this returns
    if (st == CIR_OK) {
        const hipError_t e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) st = hip_fail(e, "hipStreamSynchronize", __FILE__, __LINE__);
    }
    return st;
}
}  // namespace

extern "C" int cir_stream_push(cir_ctx* c, int first, int count, const float* xyz, const float* box) try {
    if (!c) return fail(CIR_ERR_ARG, "cir_stream_push: null context");
    const int st = stream_push(c, first, count, xyz, box);
    if (st != CIR_OK) c->streamed.reset();  // a failed push closes the stream
    return st;
} catch (...) {
    c->streamed.reset();
    return cir::host_exception("cir_stream_push");
}

extern "C" int cir_stream_end(cir_ctx* c) try {
    if (!c) return fail(CIR_ERR_ARG, "cir_stream_end: null context");
    std::unique_ptr<StreamRun, StreamDeleter> s = std::move(c->streamed);
    if (!s) return fail(CIR_ERR_ARG, "cir_stream_end: no stream is open (cir_stream_begin)");
    if (s->c0 < s->run.nu)
        return fail(CIR_ERR_FRAME, "cir_stream_end: trajectory frame " + std::to_string(s->run.S.uniq[s->c0]) +
                                     ", which the samples use, was never pushed");
    const int st = s->run.finish();
    hipEvent_t e1;
    if (hipEventCreate(&e1) == hipSuccess) {
        float ms = 0.f;
        if (hipEventRecord(e1, c->stream) == hipSuccess && hipEventSynchronize(e1) == hipSuccess &&
            hipEventElapsedTime(&ms, s->e0, e1) == hipSuccess)
            c->seconds = ms * 1e-3;
        (void)hipEventDestroy(e1);
    }
    c->prof.collect(c->phase_sec, c->phase_cnt, CIR_NPHASE);
    return st;
} catch (...) {
    return cir::host_exception("cir_stream_end");
}

extern "C" int cir_set_profiling(cir_ctx* c, int on) {
    if (!c) return fail(CIR_ERR_ARG, "null context");
    c->prof.on = on != 0;
    return CIR_OK;
}

extern "C" int cir_tuning_report(char* buf, int buflen) {
    const std::string r = tune_report();
    if (buf && buflen > 0) {
        const size_t k = std::min(r.size(), size_t(buflen - 1));
        std::memcpy(buf, r.data(), k);
        buf[k] = '\0';
    }
    return int(r.size());
}

extern "C" int cir_phase_times(cir_ctx* c, double* seconds, long long* groups, int nphase) {
    if (!c || nphase < 0 || nphase > CIR_NPHASE) return fail(CIR_ERR_ARG, "cir_phase_times: bad arguments");
    for (int i = 0; i < nphase; ++i) {
        if (seconds) seconds[i] = c->phase_sec[i];
        if (groups) groups[i] = c->phase_cnt[i];
    }
    return CIR_OK;
}

extern "C" int cir_run(cir_ctx* c, int nframes, const float* xyz, const float* box, const float* times) try {
    if (!c || !xyz || !box || !times || nframes <= 0) return fail(CIR_ERR_ARG, "cir_run: bad arguments");
    const size_t nx = size_t(nframes) * c->natoms * 3;
    for (size_t i = 0; i < nx; ++i)
        if (!std::isfinite(xyz[i])) return fail(CIR_ERR_NONFINITE, "non-finite coordinate in the trajectory");
    CIR_TRY(check_boxes(c, nframes, box));
    const size_t nbx = size_t(nframes) * box_floats(c);
    DevBuf dx, db;
    CIR_TRY(dx.alloc(nx * sizeof(float)));
    CIR_TRY(db.alloc(nbx * sizeof(float)));
    CIR_HIP(hipMemcpyAsync(dx.p, xyz, nx * sizeof(float), hipMemcpyHostToDevice, c->stream));
    CIR_HIP(hipMemcpyAsync(db.p, box, nbx * sizeof(float), hipMemcpyHostToDevice, c->stream));
    return cir_run_device(c, nframes, dx.as<float>(), db.as<float>(), times);
} catch (...) {
    return cir::host_exception("cir_run");
}

// The memory pool of cir_run_many's jobs on the current device, created on
// first use and kept until the process ends.  It holds freed memory for the
// next job (release threshold: all) until the call trims it.  Null when the
// device has no stream-ordered allocator: the jobs then use hipMalloc.
static hipMemPool_t job_pool() {
    static std::mutex m;
    static std::vector<std::pair<int, hipMemPool_t>> pools;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    std::lock_guard<std::mutex> g(m);
    for (const auto& q : pools)
        if (q.first == dev) return q.second;
    hipMemPoolProps pr{};
    pr.allocType = hipMemAllocationTypePinned;
    pr.handleTypes = hipMemHandleTypeNone;
    pr.location.type = hipMemLocationTypeDevice;
    pr.location.id = dev;
    hipMemPool_t pool = nullptr;
    uint64_t keep = UINT64_MAX;
    if (hipMemPoolCreate(&pool, &pr) != hipSuccess ||
        hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep) != hipSuccess) {
        (void)hipGetLastError();
        if (pool) (void)hipMemPoolDestroy(pool);
        pool = nullptr;
    }
    pools.emplace_back(dev, pool);
    return pool;
}

// Independent jobs (boxes of any size) at once on one GPU: each job runs on
// its own stream, issued by a pool of host threads; the largest jobs start
// first.  Kernels of different jobs overlap on the device, which fills the
// GPU when single small boxes cannot.  The jobs' buffers come from a memory
// pool in their streams' order (hipFree would wait for the other jobs' work).
extern "C" int cir_run_many(int njob, cir_ctx* const* ctxs, const int* nframes, const float* const* xyz,
                            const float* const* box, const float* const* times, int nthreads) try {
    if (njob < 0 || (njob > 0 && (!ctxs || !nframes || !xyz || !box || !times)))
        return fail(CIR_ERR_ARG, "cir_run_many: bad arguments");
    for (int j = 0; j < njob; ++j)
        for (int k = 0; k < j; ++k)
            if (ctxs[j] == ctxs[k] || !ctxs[j]) return fail(CIR_ERR_ARG, "cir_run_many: contexts must be distinct");
    if (njob == 0) return CIR_OK;
    std::vector<int> order(njob);
    for (int j = 0; j < njob; ++j) order[j] = j;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return double(ctxs[a]->N) * ctxs[a]->N * nframes[a] > double(ctxs[b]->N) * ctxs[b]->N * nframes[b];
    });
    // jobs bound to the null stream get a stream of their own for the call
    std::vector<hipStream_t> made(njob, nullptr), saved(njob, nullptr);
    for (int j = 0; j < njob; ++j) saved[j] = ctxs[j]->stream;
    for (int j = 0; j < njob; ++j) {
        if (saved[j]) continue;
        const hipError_t e = hipStreamCreateWithFlags(&made[j], hipStreamNonBlocking);
        if (e != hipSuccess) {
            for (int k = 0; k < j; ++k) {
                ctxs[k]->stream = saved[k];
                if (made[k]) (void)hipStreamDestroy(made[k]);
            }
            return hip_fail(e, "hipStreamCreateWithFlags", __FILE__, __LINE__);
        }
        ctxs[j]->stream = made[j];
    }
    const int nt = std::max(1, std::min(nthreads > 0 ? nthreads : 8, njob));
    // concurrent jobs split the device memory their chunks may take
    std::vector<int> saved_share(njob);
    for (int j = 0; j < njob; ++j) {
        saved_share[j] = ctxs[j]->mem_share;
        ctxs[j]->mem_share = nt;
    }
    std::vector<int> status(njob, CIR_OK);
    std::vector<std::string> detail(njob);
    std::atomic<int> next{0};
    hipMemPool_t pool = job_pool();
    auto worker = [&]() {
        for (int q = next++; q < njob; q = next++) {
            const int j = order[q];
            const PoolStream ps{pool, ctxs[j]->stream};
            tl_pool = pool ? &ps : nullptr;
            status[j] = cir_run(ctxs[j], nframes[j], xyz[j], box[j], times[j]);
            tl_pool = nullptr;
            if (status[j] != CIR_OK) detail[j] = cir_last_error();
        }
    };
    std::vector<std::thread> threads;
    try {
        for (int t = 0; t < nt; ++t) threads.emplace_back(worker);
    } catch (...) {  // fewer threads than asked: the ones started take all jobs
    }
    if (threads.empty()) worker();
    for (auto& t : threads) t.join();
    if (pool) {
        // the jobs' buffers are freed in their streams' order: once the
        // streams have passed the frees, the pool gives the memory back
        for (int j = 0; j < njob; ++j) (void)hipStreamSynchronize(ctxs[j]->stream);
        (void)hipMemPoolTrimTo(pool, 0);
        (void)hipGetLastError();
    }
    for (int j = 0; j < njob; ++j) {
        ctxs[j]->stream = saved[j];
        ctxs[j]->mem_share = saved_share[j];
        if (made[j]) (void)hipStreamDestroy(made[j]);
    }
    for (int j = 0; j < njob; ++j)
        if (status[j] != CIR_OK) return fail(status[j], "job " + std::to_string(j) + ": " + detail[j]);
    return CIR_OK;
} catch (...) {
    return cir::host_exception("cir_run_many");
}

extern "C" int cir_get_results(cir_ctx* c, cir_results* r) try {
    if (!c || !r) return fail(CIR_ERR_ARG, "cir_get_results: null argument");
    if (!c->have) return fail(CIR_ERR_ARG, "cir_get_results: no completed run");
    const int nt = c->p.ntcfpoints;
    double* tc[3] = {r->tcf_ir, r->tcf_vv, r->tcf_vh};
    double* ft[3] = {r->ftcf_ir, r->ftcf_vv, r->ftcf_vh};
    for (int s = 0; s < 3; ++s) {
        if (tc[s]) std::memcpy(tc[s], c->tcf[s].data(), 2 * size_t(nt) * sizeof(double));
        if (ft[s]) std::memcpy(ft[s], c->ftcf[s].data(), size_t(c->nfft) * sizeof(double));
    }
    const size_t nb = size_t(c->nomega) * sizeof(double);
    if (r->sw) std::memcpy(r->sw, c->sw.data(), nb);
    if (r->pw) std::memcpy(r->pw, c->pw.data(), nb);
    if (r->rw) std::memcpy(r->rw, c->rw.data(), nb);
    if (r->rmw) std::memcpy(r->rmw, c->rmw.data(), nb);
    r->nsteps = c->nsteps;
    r->nframes_used = c->nframes_used;
    r->hist_skipped = c->hist_skipped;
    r->seconds = c->seconds;
    return CIR_OK;
} catch (...) {
    return cir::host_exception("cir_get_results");
}

extern "C" void cir_destroy(cir_ctx* c) { delete c; }

// Per-stage C entry points over the kernels above -----------------------------

extern "C" int cir_eigh_batched(int prec, int batch, int n, int ld, void* A_d, void* V_d, double* w_d, void* work_d,
                                void* stream) {
    const int ep = (prec == CIR_PREC_FP64 || prec == CIR_PREC_BF16X3) ? prec : CIR_PREC_FP32;
    return eigh_batched(ep, batch, n, ld, A_d, V_d, w_d, work_d, as_stream(stream), nullptr);
}

extern "C" int cir_ipr_hist(int prec, int batch, int n, int ld, int nchrom_mol, float avef, float omegaStart,
                            int omegaStep, int nomega, const void* V_d, const double* w_d, const double* weight_d,
                            double* pw_d, double* rw_d, double* rmw_d, long long* skipped_d, void* stream) {
    return ipr_hist(prec == CIR_PREC_FP64 ? CIR_PREC_FP64 : CIR_PREC_FP32, batch, n, ld, nchrom_mol, avef, omegaStart,
                    omegaStep, nomega, V_d, w_d, weight_d, pw_d, rw_d, rmw_d, skipped_d, as_stream(stream));
}

extern "C" int cir_spectral_density(int prec, int batch, int n, int ld, float t1, float avef, float omegaStart,
                                    int omegaStep, int nomega, const void* V_d, const double* w_d, const void* sites_d,
                                    const double* weight_d, double* sw_d, double* work_d, void* stream) {
    return spectral_density(prec == CIR_PREC_FP64 ? CIR_PREC_FP64 : CIR_PREC_FP32, batch, n, ld, t1, avef, omegaStart,
                            omegaStep, nomega, V_d, w_d, sites_d, weight_d, sw_d, work_d, as_stream(stream));
}

extern "C" int cir_propagate(int prec, int n, int ld, int ncols, int ldy, float dt, const void* V_d,
                             const double* w_d, const void* Yin_d, void* Yout_d, void* work_d, void* stream) {
    const int rp = prec == CIR_PREC_FP64 ? CIR_PREC_FP64 : CIR_PREC_FP32;
    return propagate(rp, prec, n, ld, ncols, ldy, double(dt), V_d, w_d, Yin_d, Yout_d, work_d, as_stream(stream));
}

extern "C" size_t cir_propagate_workspace(int prec, int n, int ncols) {
    return propagate_work_elems(n, ncols) * (prec == CIR_PREC_FP64 ? sizeof(double) : sizeof(float));
}

extern "C" int cir_correlate(int prec, int n, int nslot, int ldy, const int* action_d, const int* target_d,
                             const void* sites_d, const void* Y_d, double* out_d, void* stream) {
    return correlate(prec == CIR_PREC_FP64 ? CIR_PREC_FP64 : CIR_PREC_FP32, n, nslot, ldy, action_d, target_d, sites_d,
                     Y_d, out_d, as_stream(stream));
}

extern "C" size_t cir_reunit_workspace(int nslot) { return reunit_work_doubles(nslot) * sizeof(double); }

extern "C" int cir_reunitarise(int prec, int n, int nslot, int ldy, const int* mode_d, double* G0_d, void* Y_d,
                               void* work_d, void* stream) {
    const int rp = (prec == CIR_PREC_FP64) ? CIR_PREC_FP64 : CIR_PREC_FP32;
    return reunitarise(rp, n, nslot, ldy, mode_d, G0_d, Y_d, static_cast<double*>(work_d), as_stream(stream));
}

extern "C" int cir_spectrum_c2r(int nspec, int nfft, int nx, const double* X_d, double* y_d, void* stream) {
    return spectrum_c2r(nspec, nfft, nx, X_d, y_d, as_stream(stream));
}

extern "C" int cir_gemm(int prec, int M, int N, int K, int batch, const void* A, long long a_m, long long a_k,
                        long long a_b, const void* B, long long b_k, long long b_n, long long b_b, void* C,
                        long long c_m, long long c_n, long long c_b, double alpha, double beta, void* stream) {
    GemmArgs g;
    g.M = M;
    g.N = N;
    g.K = K;
    g.batch = batch;
    g.A = A;
    g.a_m = a_m;
    g.a_k = a_k;
    g.a_b = a_b;
    g.B = B;
    g.b_k = b_k;
    g.b_n = b_n;
    g.b_b = b_b;
    g.C = C;
    g.c_m = c_m;
    g.c_n = c_n;
    g.c_b = c_b;
    g.alpha = alpha;
    g.beta = beta;
    return gemm(prec, g, as_stream(stream));
}
