// filt_params.cpp -- parameter defaults, the defines.ini reader and the
// stability guard (host code of libfiltration_hip.so).
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>

#include "../../include/filtration_hip.h"
#include "filt_guard.hpp"

namespace {

// defines.h:22-26 media table and defines.h:8-15 constants.
const double kK[2] = {6.64e-11, 7.15e-12};
const double kLambda[2] = {2.7, 2.0};
const double kSwr[2] = {0.09, 0.12};
const double kM[2] = {0.4, 0.39};
const double kPd[2] = {755, 2060};

bool parse_indexed(const std::string& key, const char* base, int* index)
{
    const size_t n = strlen(base);
    if (key.size() < n + 3 || key.compare(0, n, base) != 0 || key[n] != '[' || key.back() != ']')
        return false;
    char* end = nullptr;
    const long v = strtol(key.c_str() + n + 1, &end, 10);
    if (end != key.c_str() + key.size() - 1 || v < 0 || v >= FILT_MAX_MEDIA)
        return false;
    *index = (int)v;
    return true;
}

}  // namespace

extern "C" {

void filt_default_params(filt_params* p)
{
    if (!p)
        return;
    memset(p, 0, sizeof *p);
    for (int q = 0; q < 2; q++) {
        p->K[q] = kK[q];
        p->lambda[q] = kLambda[q];
        p->S_wr[q] = kSwr[q];
        p->m[q] = kM[q];
        p->P_d[q] = kPd[q];
    }
    p->mu1 = 0.001;
    p->mu2 = 0.0009;
    p->ro01 = 1000;
    p->ro02 = 1460;
    p->S2_gr = 0.4;
    p->P_atm = 100000;
    p->g_const = 9.8;
    p->newton_iterations = 7;
    p->src_x0 = p->src_x1 = p->src_z0 = p->src_z1 = -1;
    p->n_media = 2;
}

void filt_params_finalize(filt_params* p)
{
    if (!p || p->src_x0 >= 0)
        return;
    // main.cpp:109 / cpu.cpp:346 with C integer division
    p->src_x0 = p->NX / 2 - p->source;
    p->src_x1 = p->NX / 2 + p->source;
    p->src_z0 = p->Nz / 2 - p->source;
    p->src_z1 = p->Nz / 2 + p->source;
}

// main.cpp:601-683 semantics for well-formed lines: '#' in column 0 skips the
// line; the key is everything before the first '='; the value runs from after
// '=' to the first space (or the end of the line, newline included, which
// atof/atoi ignore); later lines override earlier ones; unknown keys are
// ignored.  Lines without '=' are skipped.
int filt_read_defines(const char* path, filt_params* p)
{
    return filt::guarded([&]() -> int {
        if (!path || !p)
            return FILT_ERR_INVALID;
        FILE* fh = fopen(path, "rt");
        if (!fh)
            return FILT_ERR_IO;
        char line[250];
        bool explicit_patch = false;
        while (fgets(line, sizeof line, fh)) {
            if (line[0] == '#')
                continue;
            const char* eq = strchr(line, '=');
            if (!eq)
                continue;
            const std::string key(line, eq - line);
            const char* v0 = eq + 1;
            const char* v1 = strchr(v0, ' ');
            const std::string val = v1 ? std::string(v0, v1 - v0) : std::string(v0);
            const char* s = val.c_str();
            int q = 0;
            // the 16 keys of the reference
            if (key == "H1") p->h1 = atof(s);
            else if (key == "H2") p->h2 = atof(s);
            else if (key == "H3") p->h3 = atof(s);
            else if (key == "TAU") p->tau = atof(s);
            else if (key == "DT") p->dt = atof(s);
            else if (key == "L_W") p->l_w = atof(s);
            else if (key == "L_N") p->l_n = atof(s);
            else if (key == "C") p->c = atof(s);
            else if (key == "BETA1") p->beta1 = atof(s);
            else if (key == "BETA2") p->beta2 = atof(s);
            else if (key == "SOURCE") p->source = atoi(s);
            else if (key == "SAVE_PLOTS") p->save_plots = atoi(s);
            else if (key == "PRINT_SCREEN") p->print_screen = atoi(s);
            else if (key == "NX") p->NX = atoi(s);
            else if (key == "NY") p->Ny = atoi(s);
            else if (key == "NZ") p->Nz = atoi(s);
            // keys defines.ini documents (defines.ini:52-100)
            else if (parse_indexed(key, "K", &q)) p->K[q] = atof(s);
            else if (parse_indexed(key, "LAMBDA", &q)) p->lambda[q] = atof(s);
            else if (parse_indexed(key, "S_WR", &q)) p->S_wr[q] = atof(s);
            else if (parse_indexed(key, "M", &q)) p->m[q] = atof(s);
            else if (parse_indexed(key, "P_D", &q)) p->P_d[q] = atof(s);
            else if (key == "MU1") p->mu1 = atof(s);
            else if (key == "MU2") p->mu2 = atof(s);
            else if (key == "RO_W") p->ro01 = atof(s);
            else if (key == "RO_N") p->ro02 = atof(s);
            else if (key == "S2_GR") p->S2_gr = atof(s);
            else if (key == "P_ATM") p->P_atm = atof(s);
            else if (key == "G_CONST") p->g_const = atof(s);
            // extensions of this backend
            else if (key == "NEWTON_ITERATIONS") p->newton_iterations = atoi(s);
            else if (key == "N_MEDIA") p->n_media = atoi(s);
            else if (key == "ROS_SCHEME") p->roS_scheme = atoi(s);
            else if (key == "SRC_X0") { p->src_x0 = atoi(s); explicit_patch = true; }
            else if (key == "SRC_X1") { p->src_x1 = atoi(s); explicit_patch = true; }
            else if (key == "SRC_Z0") { p->src_z0 = atoi(s); explicit_patch = true; }
            else if (key == "SRC_Z1") { p->src_z1 = atoi(s); explicit_patch = true; }
        }
        fclose(fh);
        if (!explicit_patch)
            p->src_x0 = p->src_x1 = p->src_z0 = p->src_z1 = -1;
        filt_params_finalize(p);
        return FILT_OK;
    });
}

// Explicit-scheme guard.  In a water-filled cell stage 3 + the Newton act as
// an explicit diffusion step for P with D = K/(mu1*phi*beta1); the
// checkerboard mode is stable for r = 4*dt*D*sum(1/h^2) <= 2.  Two-phase
// cells have D = K*(k1/mu1 + k2/mu2) / (phi*(beta1*(1-S) + beta2*S)); the
// guard also evaluates that at the source saturation S2_gr.
double filt_stability_number(const filt_params* p, const double* perm, const double* poro,
                             size_t n_cells)
{
    if (!p)
        return NAN;
    const double inv_h2 = 1.0 / (p->h1 * p->h1) + 1.0 / (p->h2 * p->h2) +
                          (p->Nz > 2 ? 1.0 / (p->h3 * p->h3) : 0.0);
    const int nm = p->n_media > 0 ? p->n_media : 1;
    double kmax = 0, phimin = INFINITY, factor = 0;
    for (int q = 0; q < nm; q++) {
        kmax = fmax(kmax, p->K[q]);
        phimin = fmin(phimin, p->m[q]);
        // mobility / storage factor relative to water-only, at S = S2_gr
        const double S = p->S2_gr, swr = p->S_wr[q], lam = p->lambda[q];
        const double se = (1. - S - swr) / (1. - swr);
        const double k1 = pow(se, (2. + 3. * lam) / lam);
        const double k2 = (1. - se) * (1. - se) * (1 - pow(se, (2. + lam) / lam));
        const double store = p->beta1 * (1 - S) + p->beta2 * S;
        const double two = (k1 / p->mu1 + k2 / p->mu2) / store;
        const double one = 1.0 / (p->mu1 * p->beta1);
        factor = fmax(factor, fmax(one, two));
    }
    double kphi = kmax / phimin;
    if (perm || poro) {
        double best = 0;
        for (size_t c = 0; c < n_cells; c++) {
            const double K = perm ? perm[c] : kmax;
            const double ph = poro ? poro[c] : phimin;
            best = fmax(best, K / ph);
        }
        kphi = best;
    }
    // Fluxes carry ro(P) while the storage term carries d(ro)/dP = ro01*beta1:
    // D grows with the density at the bottom of the initial hydrostatic column,
    // ro/ro01 = 1 + beta1*ro01*g*(Ny-1)*h2.
    const double depth = (p->Ny > 1 ? p->Ny - 1 : 0) * p->h2;
    const double rho_factor = fmax(1.0, 1.0 + p->beta1 * p->ro01 * p->g_const * depth);
    return 4.0 * p->dt * kphi * factor * rho_factor * inv_h2;
}

}  // extern "C"
