// cli.cpp -- the calcIR.exe command line (calcIR.cu:9-87, 173-227, 323-352,
// 765-1011) over the libcalcir pipeline.
//
// Usage: calcir <input-file>.  The trajectory key xtcf names the frame
// source: a GROMACS XTC file, or "synth:<nwater>:<nframes>:<seed>[:<box_nm>]"
// for a seeded synthetic rigid-water trajectory (frames 10 fs apart, from
// t = 0).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <string>
#include <vector>

#include "calcir.h"

namespace {

bool parse_synth(const char* spec, int* nwater, int* nframes, unsigned long long* seed, float* box) {
    if (std::strncmp(spec, "synth:", 6) != 0) return false;
    *box = 0.0f;
    const int got = std::sscanf(spec + 6, "%d:%d:%llu:%f", nwater, nframes, seed, box);
    if (got < 3 || *nwater <= 0 || *nframes < 2) return false;
    if (got < 4 || !(*box > 0.0f)) *box = cir_synth_box_nm(*nwater);
    return true;
}

// Exit status of a run stopped at max_frames (distinct from EXIT_FAILURE so
// that a job script can loop on it).
constexpr int kExitIncomplete = 3;

int fail_msg(int st) {
    std::printf("ERROR: %s: %s\n", cir_strerror(st), cir_last_error());
    return EXIT_FAILURE;
}

// The diagonals of full boxes, as calcIR.cu:404-405 uses them.
std::vector<float> diagonal_boxes(const std::vector<float>& box9, int nframes) {
    std::vector<float> box(size_t(nframes) * 3);
    for (int f = 0; f < nframes; ++f)
        for (int k = 0; k < 3; ++k) box[3 * size_t(f) + k] = box9[9 * size_t(f) + 4 * k];
    return box;
}

// A batched run fed from the XTC file in blocks of frames, so that the
// trajectory is never held whole (the reference reads each frame on demand,
// calcIR.cu:380-386).
int stream_xtc(cir_ctx* ctx, const cir_params& p, int natoms, int nframes, const std::vector<float>& times) {
    constexpr int kBlock = 64;
    int st = cir_stream_begin(ctx, nframes, times.data());
    if (st != CIR_OK) return st;
    cir_xtc* rd = nullptr;
    st = cir_xtc_open(p.xtcf, natoms, &rd);
    if (st != CIR_OK) return st;
    std::vector<float> xyz(size_t(kBlock) * natoms * 3), box9(size_t(kBlock) * 9);
    for (int f0 = 0; f0 < nframes && st == CIR_OK; f0 += kBlock) {
        const int k = std::min(kBlock, nframes - f0);
        st = cir_xtc_next(rd, k, xyz.data(), box9.data(), nullptr);
        if (st != CIR_OK) break;
        const std::vector<float> box = p.triclinic ? box9 : diagonal_boxes(box9, k);
        st = cir_stream_push(ctx, f0, k, xyz.data(), box.data());
    }
    cir_xtc_close(rd);
    return st == CIR_OK ? cir_stream_end(ctx) : st;
}

}  // namespace

extern "C" int cir_main(int argc, char** argv) try {
    if (argc != 2) {
        std::printf("Usage:\n"
                    "\tInclude as the first argument the name of an input file. No other arguments are allowed.\n");
        return EXIT_FAILURE;
    }
    const auto t0 = std::chrono::steady_clock::now();
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) == hipSuccess)
        std::printf("\nGPU INFO:\n\tDevice name: %s\n\tMemory: %g gb\n", prop.name, prop.totalGlobalMem / (1.E9));

    std::printf("\n>>> Setting default parameters\n");
    cir_params p;
    cir_params_default(&p);
    int st = cir_params_read(argv[1], &p, 1);
    if (st != CIR_OK) return EXIT_FAILURE;
    cir_params_echo(&p);

    int imodel = 0, ispecies = 0;
    if (cir_resolve_model(p.model, &imodel) != CIR_OK) {
        std::printf("WARNING: model: %s is not recognized. Check input file. Aborting...\n", p.model);
        return EXIT_FAILURE;
    }
    if (cir_resolve_species(p.species, &ispecies) != CIR_OK) {
        std::printf("WARNING: species: %s is not recognized. Check input file. Aborting...\n", p.species);
        return EXIT_FAILURE;
    }

    int nwater = 0, nframes = 0, natoms = 0;
    unsigned long long seed = 0;
    float boxl = 0.0f;
    std::vector<float> xyz, times, box;
    bool streamed = false;
    if (parse_synth(p.xtcf, &nwater, &nframes, &seed, &boxl)) {
        natoms = 4 * nwater;
        if (p.natom_mol != 4) {
            std::printf("WARNING: the synthetic source writes 4-site waters; natom_mol must be 4.\n");
            return EXIT_FAILURE;
        }
        xyz.resize(size_t(nframes) * natoms * 3);
        times.resize(nframes);
        if (p.triclinic) {
            std::printf("WARNING: the synthetic source writes cubic boxes; set triclinic 0.\n");
            return EXIT_FAILURE;
        }
        box.assign(size_t(nframes) * 3, boxl);
        st = cir_synth_trajectory(nwater, nframes, seed, boxl, xyz.data(), times.data());
        if (st != CIR_OK) return fail_msg(st);
    } else {
        // calcIR.cu:212-219: the trajectory must open and hold frames
        if (cir_xtc_scan(p.xtcf, &natoms, &nframes) != CIR_OK || nframes < 1) {
            std::printf("WARNING: The file %s could not be opened. Is the name correct?\n", p.xtcf);
            return EXIT_FAILURE;
        }
        if (nframes < 2) {
            std::printf("WARNING: %s holds fewer than two frames.\n", p.xtcf);
            return EXIT_FAILURE;
        }
        // the frame times and boxes first (headers only)
        times.resize(nframes);
        std::vector<float> box9(size_t(nframes) * 9);
        st = cir_xtc_times(p.xtcf, nframes, times.data(), box9.data());
        if (st != CIR_OK) return fail_msg(st);
        // calcIR.cu:342-345 rounds the frame spacing at the file's precision,
        // as read with frame 1 (files of at most 9 atoms store none: keep the
        // default)
        float xprec = -1.0f;
        st = cir_xtc_precision(p.xtcf, 1, &xprec);
        if (st != CIR_OK) return fail_msg(st);
        if (xprec > 0.0f) p.time_prec = xprec;
        // rectangular boxes go through as their diagonal (calcIR.cu:404-405);
        // any off-diagonal element switches to the triclinic minimum image
        bool tri = p.triclinic != 0;
        for (int f = 0; f < nframes && !tri; ++f)
            for (int k = 0; k < 9; ++k)
                if (k % 4 != 0 && box9[9 * size_t(f) + k] != 0.0f) tri = true;
        if (tri) {
            p.triclinic = 1;
            std::printf(">>> Triclinic box: using the triclinic minimum image.\n");
        }
        // batched runs read the coordinates block by block while they run, as
        // the reference reads frames on demand; checkpointed runs load them
        streamed = p.mode == CIR_MODE_BATCHED && p.ckpt_every == 0 && p.resume == 0 && p.max_frames == 0;
        if (!streamed) {
            xyz.resize(size_t(nframes) * natoms * 3);
            st = cir_xtc_read_full(p.xtcf, 0, nframes, natoms, xyz.data(), box9.data(), nullptr, nullptr);
            if (st != CIR_OK) return fail_msg(st);
            box = tri ? box9 : diagonal_boxes(box9, nframes);
        }
    }
    const int nmol = natoms / p.natom_mol, nchrom = nmol * p.nchrom_mol;
    std::printf(">>> Will read the trajectory from: %s.\n", p.xtcf);
    std::printf(">>> Found %d atoms and %d molecules.\n", natoms, nmol);
    std::printf(">>> Found %d chromophores.\n", nchrom);
    // calcIR.cu:345-346: the rounded spacing the schedule uses
    std::printf(">>> Frame time offset is: %f (ps)\n",
                float(std::round((times[1] - times[0]) * p.time_prec) / (1. * p.time_prec)));
    std::printf("\n>>> Now calculating the absorption spectrum\n");
    std::printf("----------------------------------------------------------\n");
    std::fflush(stdout);

    cir_ctx* ctx = nullptr;
    st = cir_create(&p, natoms, nullptr, &ctx);
    if (st != CIR_OK) return fail_msg(st);
    if (streamed)
        st = stream_xtc(ctx, p, natoms, nframes, times);
    else
        st = cir_run(ctx, nframes, xyz.data(), box.data(), times.data());
    if (st == CIR_INCOMPLETE) {
        // max_frames reached: the checkpoint holds the state; no output files
        std::printf("\n>>> %s\n>>> Run again with resume 1 to continue.\n", cir_last_error());
        cir_destroy(ctx);
        return kExitIncomplete;
    }
    if (st != CIR_OK) {
        cir_destroy(ctx);
        return fail_msg(st);
    }
    const int nt = p.ntcfpoints, nom = cir_nomega(&p), nfft = nt + p.nzeros;
    std::vector<double> tc(3 * 2 * size_t(nt)), ft(3 * size_t(nfft)), dist(4 * size_t(nom));
    cir_results r{};
    r.tcf_ir = tc.data();
    r.tcf_vv = tc.data() + 2 * nt;
    r.tcf_vh = tc.data() + 4 * nt;
    r.ftcf_ir = ft.data();
    r.ftcf_vv = ft.data() + nfft;
    r.ftcf_vh = ft.data() + 2 * size_t(nfft);
    r.sw = dist.data();
    r.pw = dist.data() + nom;
    r.rw = dist.data() + 2 * nom;
    r.rmw = dist.data() + 3 * nom;
    st = cir_get_results(ctx, &r);
    cir_destroy(ctx);
    if (st != CIR_OK) return fail_msg(st);
    if (r.hist_skipped > 0)
        std::printf("WARNING: %lld mode-steps fell outside omegaStart..omegaStop and were not binned.\n",
                    (long long)r.hist_skipped);
    std::printf("\n\n----------------------------------------------------------\n");
    std::printf("Finishing up...\n");
    st = cir_write_outputs(&p, &r);
    if (st != CIR_OK) return fail_msg(st);
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::printf("\n>>> Done with the calculation in %f seconds.\n", secs);
    std::printf(">>> %lld (sample, tcf point) steps over %lld distinct frames, %.3f s on the device.\n",
                (long long)r.nsteps, (long long)r.nframes_used, r.seconds);
    return 0;
} catch (const std::exception& e) {
    std::printf("ERROR: %s\n", e.what());
    return EXIT_FAILURE;
}
