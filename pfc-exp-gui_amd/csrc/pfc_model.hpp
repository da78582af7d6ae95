// pfc_model.hpp - host-side model constants, initial fields and spectral
// tables for the Misfit / CVD binary PFC models.
//
// Everything here runs on the host in the reference's own float/double
// expression types (all model constants are float; pow(float,int) promotes to
// double, sqrt/cos of a float stay float), so the tables the kernels consume
// carry the reference's values.
#pragma once
#include <cstdint>
#include <vector>

#include "../../include/pfc_hip.h"

namespace pfc {

struct Model {
    int model, L, M, N;
    float u0, sig, con0, dt, dx, axx, b0;
    int grainx, grainy, grainz, dnx, dny, dnz;
    unsigned int seed;
    float a_lattice, b_lattice, gaa1, gaa2, gab1, gab2, niu, ksi;
    float ka0, ka1, ka2, sigMa1, sigMa2, kb0, kb1, kb2, sigMb1, sigMb2;
    float c0, mn, mc, alpha, w;
};

// Reference constants of `model` with the CLI-visible values of `p`.
Model make_model(const pfc_params& p);
void default_params(pfc_params& p, int model);

// k tables krefl/krefm/krefn (misfit.cpp:455-462).
void k_tables(const Model& m, std::vector<float>& kl, std::vector<float>& km, std::vector<float>& kn);

// C2A/C2B on the folded half spectrum: out[((il*K + kc)*Nh + jn)*2 + {0,1}],
// il in [0, L/2], kc in [0, M/2], jn in [0, N/2], K = M/2+1, Nh = N/2+1
// (misfit.cpp:700-750).  Also C2A(0), C2B(0).
void c2_tables_folded(const Model& m, std::vector<float>& out, float* c2a0, float* c2b0);

// Reference initialisers; phi/con: L*M*N doubles.
void init_fields(const Model& m, double* phi, double* con);

uint64_t splitmix64_next(uint64_t& state);
double rng_u01(uint64_t& state);

}  // namespace pfc
