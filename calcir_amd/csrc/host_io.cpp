// host_io.cpp -- the drop-in host boundary of calcIR: defaults, input-file
// parser, parameter echo, model/species resolution, frame selection, the 13
// output writers and the calcIR.exe command line.
//
// The parser deliberately reproduces the reference's fscanf-based loop
// (calcIR.cu:1556): "key value" pairs, the rest of the line ignored, a line
// holding a single token takes the next line's first token as its value,
// and '#key' lines are unknown keys (warned, ignored).

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"

namespace {

void copy_str(char* dst, const char* src) {
    std::strncpy(dst, src, CIR_STRLEN - 1);
    dst[CIR_STRLEN - 1] = '\0';
}

}  // namespace

extern "C" void cir_params_default(cir_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    // calcIR.cu:36-59
    copy_str(p->xtcf, "traj.xtc");
    copy_str(p->outf, "spec");
    copy_str(p->cptf, "spec");
    copy_str(p->model, "e3b3");
    copy_str(p->species, " ");
    p->imap = 0;
    p->ntcfpoints = 150;
    p->nsamples = 1;
    p->sampleEvery = 10.0f;
    p->omegaStart = 2000;
    p->omegaStop = 5000;
    p->omegaStep = 5;
    p->natom_mol = 4;
    p->nchrom_mol = 2;
    p->nzeros = 25600;
    p->dt = 0.010f;
    p->beginTime = 0;
    p->t1 = 0.260f;
    p->avef = 3415.2f;
    // extensions
    p->cutoff = 0.7831f;  // calcIR.cu:1035
    p->precision = CIR_PREC_FP32;
    p->mode = CIR_MODE_BATCHED;
    p->chunk = 0;
    p->ckpt_every = 0;
    p->resume = 0;
    p->max_frames = 0;
    p->adj_store = 0;
    p->triclinic = 0;
    p->reunit_every = 0;
    p->origin_stride = 0;
    p->oh_per_molecule = 0;
    p->msite_bisector = 0;
    p->time_prec = 1000.0f;
    p->deterministic = 0;
}

extern "C" int cir_params_read(const char* path, cir_params* p, int verbose) try {
    if (!path || !p) return cir::fail(CIR_ERR_ARG, "cir_params_read: null argument");
    FILE* f = std::fopen(path, "r");
    if (!f) {
        if (verbose)
            std::printf("ERROR: Could not open %s. The first argument should contain  a  vaild\nfile name that points "
                        "to a file containing the simulation parameters.",
                        path);
        return cir::fail(CIR_ERR_IO, std::string("could not open input file ") + path);
    }
    if (verbose) std::printf(">>> Reading parameters from input file %s\n", path);
    char para[1024] = {0}, value[1024] = {0};
    // calcIR.cu:1556-1638
    while (std::fscanf(f, "%1023s%1023s%*[^\n]", para, value) != EOF) {
        if (!std::strcmp(para, "xtcf")) std::sscanf(value, "%255s", p->xtcf);
        else if (!std::strcmp(para, "outf")) std::sscanf(value, "%255s", p->outf);
        else if (!std::strcmp(para, "cptf")) std::sscanf(value, "%255s", p->cptf);
        else if (!std::strcmp(para, "model")) std::sscanf(value, "%255s", p->model);
        else if (!std::strcmp(para, "ntcfpoints")) std::sscanf(value, "%d", &p->ntcfpoints);
        else if (!std::strcmp(para, "nsamples")) std::sscanf(value, "%d", &p->nsamples);
        else if (!std::strcmp(para, "sampleEvery")) std::sscanf(value, "%f", &p->sampleEvery);
        else if (!std::strcmp(para, "omegaStep")) std::sscanf(value, "%d", &p->omegaStep);
        else if (!std::strcmp(para, "natom_mol")) std::sscanf(value, "%d", &p->natom_mol);
        else if (!std::strcmp(para, "nchrom_mol")) std::sscanf(value, "%d", &p->nchrom_mol);
        else if (!std::strcmp(para, "nzeros")) std::sscanf(value, "%d", &p->nzeros);
        else if (!std::strcmp(para, "map")) std::sscanf(value, "%d", &p->imap);
        else if (!std::strcmp(para, "species")) std::sscanf(value, "%255s", p->species);
        else if (!std::strcmp(para, "dt")) std::sscanf(value, "%f", &p->dt);
        else if (!std::strcmp(para, "t1")) std::sscanf(value, "%f", &p->t1);
        else if (!std::strcmp(para, "avef")) std::sscanf(value, "%f", &p->avef);
        else if (!std::strcmp(para, "beginTime")) std::sscanf(value, "%f", &p->beginTime);
        else if (!std::strcmp(para, "omegaStart")) std::sscanf(value, "%f", &p->omegaStart);
        else if (!std::strcmp(para, "omegaStop")) std::sscanf(value, "%f", &p->omegaStop);
        // extensions
        else if (!std::strcmp(para, "cutoff")) std::sscanf(value, "%f", &p->cutoff);
        else if (!std::strcmp(para, "chunk")) std::sscanf(value, "%d", &p->chunk);
        else if (!std::strcmp(para, "checkpoint_every")) std::sscanf(value, "%d", &p->ckpt_every);
        else if (!std::strcmp(para, "resume")) std::sscanf(value, "%d", &p->resume);
        else if (!std::strcmp(para, "max_frames")) std::sscanf(value, "%d", &p->max_frames);
        else if (!std::strcmp(para, "adjoint_store")) std::sscanf(value, "%d", &p->adj_store);
        else if (!std::strcmp(para, "triclinic")) std::sscanf(value, "%d", &p->triclinic);
        else if (!std::strcmp(para, "reunitarise_every")) std::sscanf(value, "%d", &p->reunit_every);
        else if (!std::strcmp(para, "origin_stride_frames")) std::sscanf(value, "%d", &p->origin_stride);
        else if (!std::strcmp(para, "oh_per_molecule")) std::sscanf(value, "%d", &p->oh_per_molecule);
        else if (!std::strcmp(para, "msite_bisector")) std::sscanf(value, "%d", &p->msite_bisector);
        else if (!std::strcmp(para, "time_precision")) std::sscanf(value, "%f", &p->time_prec);
        else if (!std::strcmp(para, "deterministic")) std::sscanf(value, "%d", &p->deterministic);
        else if (!std::strcmp(para, "precision")) {
            if (!std::strcmp(value, "fp64")) p->precision = CIR_PREC_FP64;
            else if (!std::strcmp(value, "bf16x3")) p->precision = CIR_PREC_BF16X3;
            else if (!std::strcmp(value, "fp32")) p->precision = CIR_PREC_FP32;
            else if (verbose) std::printf("WARNING: precision %s not recognized, keeping the current one.\n", value);
        } else if (!std::strcmp(para, "mode")) {
            if (!std::strcmp(value, "plain")) p->mode = CIR_MODE_PLAIN;
            else if (!std::strcmp(value, "batched")) p->mode = CIR_MODE_BATCHED;
            else if (verbose) std::printf("WARNING: mode %s not recognized, keeping the current one.\n", value);
        } else if (verbose) {
            std::printf("WARNING: Parameter %s in input file %s not recognized, ignoring.\n", para, path);
        }
    }
    std::fclose(f);
    if (verbose) std::printf(">>> Done reading input file and setting parameters\n");
    return CIR_OK;
} catch (...) {
    return cir::host_exception("cir_params_read");
}

extern "C" void cir_params_echo(const cir_params* p) {
    // calcIR.cu:70-87
    std::printf("\tSetting xtc file %s\n", p->xtcf);
    std::printf("\tSetting default file name to %s\n", p->outf);
    std::printf("\tSetting model to %s\n", p->model);
    std::printf("\tSetting the number of tcf points to %d\n", p->ntcfpoints);
    std::printf("\tSetting nsamples to %d\n", p->nsamples);
    std::printf("\tSetting sampleEvery to %f (ps)\n", p->sampleEvery);
    std::printf("\tSetting omegaStep to %d\n", p->omegaStep);
    std::printf("\tSetting natom_mol to %d\n", p->natom_mol);
    std::printf("\tSetting nchrom_mol to %d\n", p->nchrom_mol);
    std::printf("\tSetting nzeros to %d\n", p->nzeros);
    std::printf("\tSetting map to %d\n", p->imap);
    std::printf("\tSetting species to %s\n", p->species);
    std::printf("\tSetting omegaStart to %f\n", p->omegaStart);
    std::printf("\tSetting omegaStop to %f\n", p->omegaStop);
    std::printf("\tSetting dt to %f\n", p->dt);
    std::printf("\tSetting t1 to %f (ps)\n", p->t1);
    std::printf("\tSetting avef to %f\n", p->avef);
    std::printf("\tSetting equilibration time to %f (ps)\n", p->beginTime);
    // extension keys are echoed only when set, so a reference input echoes as before
    if (p->ckpt_every > 0 || p->resume != 0 || p->max_frames > 0)
        std::printf("\tCheckpoint %s_ckpt.bin: every %d frames, resume %d, max_frames %d\n", p->cptf, p->ckpt_every,
                    p->resume, p->max_frames);
    if (p->origin_stride > 0) std::printf("\tTime origins every %d frames\n", p->origin_stride);
    if (p->oh_per_molecule) std::printf("\tChromophore k of a molecule is its O-H%s stretch\n", "(1+k)");
    if (p->msite_bisector) std::printf("\tThree-site water: -1.04 site on the HOH bisector\n");
    if (p->deterministic) std::printf("\tDeterministic reductions\n");
}

extern "C" int cir_resolve_model(const char* model, int* imodel) try {
    // calcIR.cu:325-330
    if (!model || !imodel) return cir::fail(CIR_ERR_ARG, "null argument");
    if (!std::strcmp(model, "tip4p2005") || !std::strcmp(model, "e3b3")) *imodel = 1;
    else if (!std::strcmp(model, "tip4p") || !std::strcmp(model, "e3b2")) *imodel = 0;
    else return cir::fail(CIR_ERR_MODEL, std::string("model ") + model + " is not recognized");
    return CIR_OK;
} catch (...) {
    return cir::host_exception("cir_resolve_model");
}

extern "C" int cir_resolve_species(const char* species, int* ispecies) try {
    // calcIR.cu:332-339
    if (!species || !ispecies) return cir::fail(CIR_ERR_ARG, "null argument");
    if (!std::strcmp(species, "H2O")) *ispecies = 0;
    else if (!std::strcmp(species, "HOD/D2O")) *ispecies = 1;
    else if (!std::strcmp(species, "HOD/H2O")) *ispecies = 2;
    else if (!std::strcmp(species, "D2O")) *ispecies = 3;
    else return cir::fail(CIR_ERR_SPECIES, std::string("species ") + species + " is not recognized");
    return CIR_OK;
} catch (...) {
    return cir::host_exception("cir_resolve_species");
}

extern "C" int cir_nomega(const cir_params* p) {
    // calcIR.cu:94, float arithmetic then truncation
    return int((p->omegaStop - p->omegaStart) / float(p->omegaStep) + 1);
}

extern "C" float cir_minimage(float dx, float boxl) {
    // calcIR.cu:1457-1460
    return dx - boxl * std::round(dx / boxl);
}

extern "C" int cir_frame_index(const cir_params* p, int sample, int point, float frame_dt, float* desired_time) {
    // calcIR.cu:378-379, float arithmetic as in the reference; origin_stride
    // gives sampleEvery in frames
    const float every = (p->origin_stride > 0) ? float(p->origin_stride) * frame_dt : p->sampleEvery;
    const float t = float(sample) * every + p->beginTime + p->dt * float(point);
    if (desired_time) *desired_time = t;
    return int(std::round(t / frame_dt));
}

namespace {

int write_pairs(const std::string& name, const std::vector<std::pair<float, float>>& rows, bool trailing_space) {
    FILE* f = std::fopen(name.c_str(), "w");
    if (!f) return cir::fail(CIR_ERR_IO, "could not open " + name + " for writing");
    for (const auto& r : rows) {
        if (trailing_space)
            std::fprintf(f, "%g %g \n", r.first, r.second);
        else
            std::fprintf(f, "%g %g\n", r.first, r.second);
    }
    std::fclose(f);
    return CIR_OK;
}

}  // namespace

extern "C" int cir_write_outputs(const cir_params* p, const cir_results* r) try {
    if (!p || !r) return cir::fail(CIR_ERR_ARG, "cir_write_outputs: null argument");
    const int ntcf = p->ntcfpoints, nz = p->nzeros, nfft = ntcf + nz;
    const int nomega = cir_nomega(p);
    const std::string out(p->outf);
    std::vector<std::pair<float, float>> rows;
    // TCFs, calcIR.cu:843-870
    const double* tc[3] = {r->tcf_ir, r->tcf_vv, r->tcf_vh};
    const char* names[3][2] = {{"_irrtcf.dat", "_iritcf.dat"}, {"_vvrtcf.dat", "_vvitcf.dat"}, {"_vhrtcf.dat", "_vhitcf.dat"}};
    for (int s = 0; s < 3; ++s) {
        if (!tc[s]) continue;
        for (int part = 0; part < 2; ++part) {
            rows.clear();
            for (int i = 0; i < ntcf; ++i) rows.emplace_back(float(i) * p->dt, float(tc[s][2 * i + part]));
            CIR_TRY(write_pairs(out + names[s][part], rows, true));
        }
    }
    // spectral density and distributions, calcIR.cu:875-890
    const double* dist[4] = {r->sw, r->pw, r->rw, r->rmw};
    const char* dnames[4] = {"_spdn.dat", "_Pw.dat", "_Rw.dat", "_Rmw.dat"};
    for (int s = 0; s < 4; ++s) {
        if (!dist[s]) continue;
        rows.clear();
        for (int i = 0; i < nomega; ++i)
            rows.emplace_back(float(p->omegaStart + float(p->omegaStep * i)), float(dist[s][i]));
        CIR_TRY(write_pairs(out + dnames[s], rows, false));
    }
    // lineshapes, calcIR.cu:893-919 (float arithmetic, descending frequency;
    // the second loop bound keeps the reference's operator precedence)
    const double* ft[3] = {r->ftcf_ir, r->ftcf_vv, r->ftcf_vh};
    const char* lnames[3] = {"_irls.dat", "_vvls.dat", "_vhls.dat"};
    const float factor = float(2 * cir::kPi * cir::kHbar / double(p->dt * float(nfft)));
    for (int s = 0; s < 3; ++s) {
        if (!ft[s]) continue;
        rows.clear();
        for (int i = nfft / 2; i < nfft; ++i) {
            const float freq = float(-1 * (i - ntcf - nz)) * factor + p->avef;
            if (freq <= p->omegaStop) rows.emplace_back(freq, float(ft[s][i]) / (factor * float(nfft)));
        }
        for (int i = 0; i < ntcf + nz / 2; ++i) {
            const float freq = float(-1 * i) * factor + p->avef;
            if (freq >= p->omegaStart) rows.emplace_back(freq, float(ft[s][i]) / (factor * float(nfft)));
        }
        CIR_TRY(write_pairs(out + lnames[s], rows, false));
    }
    return CIR_OK;
} catch (...) {
    return cir::host_exception("cir_write_outputs");
}
