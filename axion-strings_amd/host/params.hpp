// params.hpp -- run parameters of the host driver.
//
// `struct parameters` keeps every field of the reference's struct
// (src/common/common.hpp:30-69) with the same names and types, so a
// reference parameter file parses unchanged; the fields after the marked
// line are new and optional (defaults below), which keeps default.param
// valid.
#pragma once

#include <stdio.h>

#define AXS_MAX_LEN 200

struct parameters {
    // ---- the reference's fields (common.hpp:35-68) ----
    float lambda;
    int enable_PRS;
    float T_initial;
    int NDIMS;
    int N;
    float space_step;
    float time_step;
    int stencil_setting;
    unsigned int seed;

    int write_output_file;
    char output_file_path[AXS_MAX_LEN];

    int save_snapshots;
    int n_snapshots;
    char output_directory[AXS_MAX_LEN];
    int save_fields;
    int save_strings;
    int save_pk;

    int sample_time_series;
    int n_samples;
    char ts_output_path[AXS_MAX_LEN];

    int sample_strings;
    int sample_background;

    int thr;

    int enable_amr;
    float refinement_threshold;
    int init_from_snapshot;
    float tau_initial;

    // ---- optional additions (absent from a reference file = default) ----
    int n_steps;            // 0: the reference's light-crossing count (s_run.cpp:82)
    char precision[16];     // "fp64" (plain form) or "fp32"
    char init_mode[16];     // "thermal" (init.cpp), "kinv" (k^-1 field)
    int init_on_device;     // 1: generate the initial field in device memory
    int sample_energy;      // 1: append E_kin, E_grad, E_pot to the time series
    char pk_binning[16];    // "linear" (pk_estimation.cpp) or "log"
    int pk_nbins;           // bins of the "log" form
    int save_full_fields;   // 1: 3D field snapshots hold all N^3 values
    int device;             // HIP device ordinal
    char checkpoint_path[AXS_MAX_LEN];  // "" = none; full state written at the end
    char restart_path[AXS_MAX_LEN];     // "" = none; start from this checkpoint
    char eom[24];           // "field_rescaled" (s_run.cpp:215) or "hamiltonian" (:216)
    char laplacian[16];     // "stencil" (s_integrate.cpp:322) or "spectral" (:326-328)
    int debug_nan;          // 1: the NaN scan before every step (s_run.cpp:4-12, 213)
};

// Gadget2-style reader with the reference's rules (read_parameters.cpp:
// 182-238): every reference tag exactly once, unknown or repeated tags are
// an error, lines starting with % are comments.  Optional tags may appear
// at most once.  Returns 0, or 1 after printing the reference's message.
int read_parameter_file(const char *fname, parameters *p);

// defaults of the optional fields
void parameters_set_defaults(parameters *p);
