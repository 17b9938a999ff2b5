// Internal interface of the reference's front ends (hh_conf.cpp): the .conf driver and the
// reference-named C API (hh_refapi.cpp) share the calculations, the CF/SF file formats and the
// surface checks.
#pragma once

#include <dlfcn.h>

#include <string>
#include <vector>

#include "../../include/hawaii_conf.h"

namespace hh {
namespace conf {

// A CF, SF or spectrum as the reference's files hold it (src/hawaii.c:4150-4264)
struct Series {
    bool sf = false;
    bool spectrum = false;      // a Spectrum (ALPHA's output, cm^-1 Amagat^-2): x in cm^-1
    std::vector<double> x, y;   // t or nu; normalised values
    double T = 0.0, ntraj = 0.0;
};

// Host surfaces of the reference's plugin ABI: the .conf's plugins (dlopen'ed by load_surfaces)
// or the functions a program hands to the reference-named API.  Null: the built-in surface.
struct HostSurfaces {
    hh_pes_fn pes = nullptr;
    hh_dpes_fn dpes = nullptr;
    hh_dipole_fn dipole = nullptr;
    int pes_system = -1, dipole_system = -1;   // hh_system_id() a plugin declares (-1: none)
    std::string pes_name = "the potential", dipole_name = "the dipole";
    void* so[2] = {nullptr, nullptr};          // plugin handles (closed with the object)
    HostSurfaces() = default;
    HostSurfaces(const HostSurfaces&) = delete;
    HostSurfaces& operator=(const HostSurfaces&) = delete;
    ~HostSurfaces()
    {
        if (so[1])
            if (void (*fin)(void) = reinterpret_cast<void (*)(void)>(dlsym(so[1], "dipole_free"))) fin();
        for (void* h : so)
            if (h) dlclose(h);
    }
};

// What a calculation reads from and returns to a caller other than the .conf driver
struct RunIO {
    long skip_draws = 0;                 // the seeded stream continues after this many draws
    long draws_used = 0;                 // out: the stream position after the call
    Series* cf = nullptr;                // out: CORRELATION_SINGLE's final CF
    std::vector<Series>* cfs = nullptr;  // out: CORRELATION_ARRAY's CFs, one per satellite temperature
    Series* sf = nullptr;                // out: PR_MU's final SF
};

int load_surfaces(const hh_conf* c, HostSurfaces& hs);
int check_surfaces(hh_engine* e, const hh_system& sys, double Rmin, double Rmax, const HostSurfaces& hs);
int system_of(const hh_conf* c, hh_system* sys);
int run_correlation_single(hh_engine* e, const hh_conf* c, const hh_conf_run_opts& o, const HostSurfaces& hs,
                           RunIO* io);
int run_correlation_array(hh_engine* e, const hh_conf* c, const hh_conf_run_opts& o, const HostSurfaces& hs,
                          RunIO* io);
int run_prmu(hh_engine* e, const hh_conf* c, const hh_conf_run_opts& o, const HostSurfaces& hs, RunIO* io);
int write_series(const char* path, const Series& s);
int read_series(const char* path, bool sf, Series& s);
int cf_to_sf(hh_engine* e, const Series& cf, Series& sf);

}  // namespace conf
}  // namespace hh
