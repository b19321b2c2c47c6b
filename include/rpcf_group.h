/*
 * rpcf_group.h -- groups of simulations of one grid shape advanced by grouped launches
 * (librpcf_hip.so only).  No counterpart in the reference, which runs one simulation per
 * process (source/main_loop.cpp:15-115); this is the layer under ensembles and the
 * serving loop: every kernel of a step is launched once for all listed members.
 *
 * A group is an rpcf_ctx created with rpcf_group_create and `capacity` slots.  Each slot
 * holds one simulation (member) with its own physics (Re, Ro, dt, Lx, Lz, base profile,
 * flow control: any rpcf_params with the group's mx, my, mz), its own operators, state and
 * AB2 history.  Slot arrays use the conventions of rpcf.h.  The single-simulation entry
 * points of rpcf.h act on slot 0.
 *
 * Results are those of the same simulation run alone: members never read each other's
 * memory, and every kernel computes a member's values with the same arithmetic whatever
 * else is in the launch.
 */
#ifndef RPCF_GROUP_H
#define RPCF_GROUP_H

#include <stddef.h>

#include "rpcf.h"

#ifdef __cplusplus
extern "C" {
#endif

/* A context of `capacity` empty slots of the grid shape of `shape` (mx, my, mz; its
 * physics fields are not used).  All device memory of the slots is allocated here. */
int rpcf_group_create(const rpcf_params* shape, int capacity, rpcf_ctx** out);
int rpcf_group_capacity(const rpcf_ctx* group);
/* Device bytes one slot of this grid shape occupies (admission by free HBM). */
int rpcf_slot_bytes(const rpcf_params* shape, size_t* bytes);

/* Put a member into a slot: its base profile, operators and the inverses of its
 * Crank-Nicolson matrices are built (initAuxMatrix, source/solver.cpp:260-320); state and
 * history are cleared.  rpcf_member_release empties the slot.  The inverses depend on
 * (n_pi_x, n_pi_y, Re, dt, u_profile) only, not on Ro, flow_control or the state: when a
 * configured slot of the group has the same values (bit for bit), its inverses are copied
 * instead of built (the matrix cache; a rotation-number sweep builds them once). */
int rpcf_member_init(rpcf_ctx* group, int slot, const rpcf_params* prm);
int rpcf_member_release(rpcf_ctx* group, int slot);
/* Number of member initialisations served by the matrix cache since the group was created. */
int rpcf_group_cache_hits(const rpcf_ctx* group, long* hits);
/* rpcf_set_state / rpcf_get_state of one slot. */
int rpcf_member_set_state(rpcf_ctx* group, int slot, const double* v_hat, const double* eta_hat,
                          long pencil_stride);
int rpcf_member_get_state(rpcf_ctx* group, int slot, double* v_hat, double* eta_hat, long pencil_stride);

/* rpcf_bootstrap / rpcf_step of the n listed slots together (each slot at most once).
 * The velocity fields of the listed members are formed from their states. */
int rpcf_group_bootstrap(rpcf_ctx* group, const int* slots, int n);
int rpcf_group_step(rpcf_ctx* group, const int* slots, int n, int nsteps);
/* Move the member in slot src_slot of `src` (another group of the same grid shape and
 * precision) into `slot` of `group`: device-to-device copies of its operators, inverses,
 * state, velocity fields and AB2 history on the group's stream; the source slot is left as
 * it is.  A serving loop builds a newcomer (rpcf_member_init, rpcf_member_set_state,
 * rpcf_group_bootstrap) in a one-slot staging group on that group's own stream while the
 * running group steps, and adopts it between two steps, so the running members wait only
 * for the copy.  Waits for both groups' streams. */
int rpcf_member_adopt(rpcf_ctx* group, int slot, rpcf_ctx* src, int src_slot);
/* enable = 1: rpcf_group_step replays one captured HIP graph of a grouped step per step
 * (captured again whenever the slot list changes; not while rpcf_timing is on). */
int rpcf_group_graphs(rpcf_ctx* group, int enable);

#ifdef __cplusplus
}
#endif

#endif /* RPCF_GROUP_H */
