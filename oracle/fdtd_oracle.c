/*
 * fdtd_oracle.c — CPU oracle for the acoustic FDTD engine.
 *
 * TEST INFRASTRUCTURE ONLY: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load this library, as the checker.  The
 * product (libafdtd.so) never links, calls or falls back to it.
 *
 * A plain-C restatement of the reference's uniform FDTD
 * (csrc/kernels/acoustic_fdtd.cpp), instantiated twice:
 *   *_f32 — float fields and arithmetic, expression order of the reference,
 *           compiled without FMA contraction (bitwise target for the FP32
 *           HIP kernels);
 *   *_f64 — the same scheme in double (the FP64 accuracy oracle).
 * Each function cites the reference lines it follows; the extension terms
 * (order-2 stencil, impedance walls, heterogeneous media, point sources,
 * receiver recording) are marked as such.
 *
 * The reference itself cannot be compiled here (it includes <Eigen/Core>,
 * which the image lacks), so this restatement is pinned by closed forms,
 * invariants and hand-derived known answers in tests/test_oracle_*.py.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define SK_EXPORT __attribute__((visibility("default")))

typedef struct OrcConfig {
    int   Nx, Ny, Nz;
    float dx, c, rho_air;
    int   plate_iz;
    float plate_mass, plate_D, plate_alpha_M, plate_beta_K;
    int   n_pml;
    int   order;      /* 4 (reference) or 2 (Yee)                        */
    float courant;    /* <= 0: reference dt rule                           */
    int   boundary;   /* 0 rigid/PML, 1 impedance (order 2 only)           */
    float refl[6];    /* x-, x+, y-, y+, z-, z+                             */
} OrcConfig;

SK_EXPORT int orc_config_size(void) { return (int)sizeof(OrcConfig); }

#define REAL float
#define SUF(x) x##_f32
#include "fdtd_oracle_impl.h"
#undef REAL
#undef SUF

#define REAL double
#define SUF(x) x##_f64
#include "fdtd_oracle_impl.h"
#undef REAL
#undef SUF
