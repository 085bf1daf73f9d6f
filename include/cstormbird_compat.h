/*
 * cstormbird_compat.h — the reference's own C interface to CompleteSailModel
 * (interfaces/cstormbird/src/lifting_line/complete_sail_model.rs:17-109,
 * src/wind.rs:7-33, src/results.rs:3-19), served by libstormbird_amd.so with
 * the simulation on the GPU (bit-exact arithmetic mode, device 0).
 *
 * Same names, struct layouts, argument meaning and return codes, so a program
 * built against cstormbird's generated header links against this library
 * unchanged.  The model's methods beyond these four are sb_csm_* in
 * stormbird_gpu.h; an error message, where the reference gives none, is in
 * sb_last_error().
 */
#ifndef CSTORMBIRD_H
#define CSTORMBIRD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Opaque model (complete_sail_model.rs:17-20). */
typedef struct CompleteSailModel CompleteSailModel;

/* wind.rs:7-12: a power-law wind profile with reference height 10 m and
 * power factor 1/9 (wind.rs:14-33). */
typedef struct WindCondition {
    double velocity;
    double direction_coming_from;
} WindCondition;

/* results.rs:3-9: per sail, total integrated force and moment and input power. */
typedef struct SingleSailResult {
    double force[3];
    double moment[3];
    double input_power;
} SingleSailResult;

/* complete_sail_model.rs:29-43: NULL on a NULL or invalid setup. */
CompleteSailModel* complete_sail_model_new(const char* setup_string);
/* complete_sail_model.rs:46-57: -1 for NULL. */
int32_t complete_sail_model_get_number_of_sails(CompleteSailModel* sail_model);
/* complete_sail_model.rs:59-66 (NULL is ignored). */
void complete_sail_model_drop(CompleteSailModel* sail_model);
/* complete_sail_model.rs:68-109: -1 for NULL pointers, -2 for zero length,
 * else the number of results written (min(sails, results_out_length)).  A
 * failure the reference would panic on returns a negative code instead. */
int32_t complete_sail_model_simulate_steady_state_condition(CompleteSailModel* sail_model,
                                                            WindCondition wind_condition, double ship_velocity,
                                                            double controller_loading,
                                                            SingleSailResult* results_out,
                                                            size_t results_out_length);

#ifdef __cplusplus
}
#endif
#endif
