// common.cpp -- error plumbing and the spectroscopic map table.
#include "common.hpp"

#include <cstring>
#include <exception>
#include <new>
#include <string>

namespace cir {

namespace {
thread_local std::string g_last_error;
}

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int status, const std::string& msg) {
    g_last_error = msg;
    return status;
}

int hip_fail(hipError_t e, const char* what, const char* file, int line) {
    std::string m = std::string(hipGetErrorString(e)) + " in " + what + " (" + file + ":" + std::to_string(line) + ")";
    g_last_error = m;
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return CIR_ERR_NOMEM;
    if (e == hipErrorNoDevice) return CIR_ERR_NODEVICE;
    return CIR_ERR_HIP;
}

int host_exception(const char* where) noexcept {
    try {
        try {
            throw;
        } catch (const std::bad_alloc&) {
            return fail(CIR_ERR_NOMEM, std::string(where) + ": host allocation failed");
        } catch (const std::exception& e) {
            return fail(CIR_ERR_ARG, std::string(where) + ": " + e.what());
        } catch (...) {
            return fail(CIR_ERR_ARG, std::string(where) + ": unknown exception");
        }
    } catch (...) {  // the detail text itself could not be allocated
        return CIR_ERR_NOMEM;
    }
}

// Map coefficients, calcIR.cu:1203-1242: imap 0 Gruenbaum 2013, imap 1 Li &
// Skinner 2010; species 0/1 (H2O, HOD/D2O) use the OH maps, 2/3 (HOD/H2O,
// D2O) the OD maps.
bool map_coef(int imap, int ispecies, MapCoef* c) {
    if (ispecies < 0 || ispecies > 3 || !c) return false;
    const bool oh = (ispecies == 0 || ispecies == 1);
    if (imap == 0) {
        if (oh)
            *c = MapCoef{{3670.2, -3541.7, -152677.0}, {0.19285, -1.7261e-5}, {1.6466, 5.7692e-4},
                         {0.1646, 11.39, 63.41}, {-1361.0, 27165.0, -1.887}};
        else
            *c = MapCoef{{2767.8, -2630.3, -102601.0}, {0.16593, -2.0632e-5}, {2.0475, 8.9108e-4},
                         {0.1646, 11.39, 63.41}, {-1361.0, 27165.0, -1.887}};
        return true;
    }
    if (imap == 1) {
        if (oh)
            *c = MapCoef{{3732.9, -3519.8, -153520.0}, {0.19318, -1.7248e-5}, {1.6102, 5.8697e-4},
                         {0.1622, 10.381, 137.6}, {-1360.8, 27171.0, -1.887}};
        else
            *c = MapCoef{{2748.2, -2572.2, -102980.0}, {0.16598, -2.0752e-5}, {1.9813, 9.1419e-4},
                         {0.1622, 10.381, 137.6}, {-1360.8, 27171.0, -1.887}};
        return true;
    }
    return false;
}

}  // namespace cir

extern "C" const char* cir_last_error(void) { return cir::g_last_error.c_str(); }

extern "C" const char* cir_strerror(int status) {
    switch (status) {
        case CIR_OK: return "success";
        case CIR_ERR_ARG: return "invalid argument";
        case CIR_ERR_IO: return "file could not be opened, read or written";
        case CIR_ERR_MODEL: return "water model not recognized";
        case CIR_ERR_SPECIES: return "species not recognized";
        case CIR_ERR_MAP: return "spectroscopic map must be 0 or 1";
        case CIR_ERR_FRAME: return "desired trajectory frame not found";
        case CIR_ERR_HIP: return "HIP runtime error";
        case CIR_ERR_NOMEM: return "out of memory";
        case CIR_ERR_EIGH: return "eigensolver failed";
        case CIR_ERR_NONFINITE: return "non-finite coordinate or box";
        case CIR_ERR_NODEVICE: return "no HIP device available";
        case CIR_INCOMPLETE: return "stopped at max_frames; resume from the checkpoint";
        default: return "unknown status";
    }
}
