/* reef3d_serve.h - serving loop over the batched hot path (libreef3d_hip.so).
 *
 * No counterpart in the reference, which runs one case per process to completion.  A
 * server takes tank requests (one box domain each: grid, initial phi/u/v/w/press, time
 * step and number of momentum_FC3 steps) from any host thread and serves them with
 * continuous batching: a worker thread admits waiting requests in arrival order while
 * their device bytes fit in the HBM budget, keeps their state on the device, advances
 * every live tank by one step per round in the same launches, and retires each tank when
 * its steps are done (its fields are copied back and its device memory is reused).  A
 * request's result is the same served alone or among others, given the same plan table
 * and controls.
 *
 * All requests share the server's physical parameters (reef_params except dt) and
 * solver settings; each request has its own grid and dt.  Times are seconds since the
 * server was created (steady clock).  Field arrays are REEF_LAYOUT_IFAST.
 */
#ifndef REEF3D_SERVE_H
#define REEF3D_SERVE_H

#include "reef3d_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct reef_server reef_server;

typedef struct reef_server_config
{
    double hbm_budget_bytes;    /* device bytes the live tanks may hold (0 = no limit) */
    int max_live;               /* tanks advanced together at most (0 = no limit) */
    int precision;              /* reef_solver_precision of every projection */
    double delta;               /* reliable-update threshold of the mixed solve, (0, 1) */
    int graphs;                 /* 1: replay each step from hipGraphs (see "graphs") */
    int F49;                    /* reinitialisation discretisation, 1 or 0 */
    reef_params shared;         /* physical and solver parameters; dt is per request */
    const int* plans;           /* nplans tile-plan entries of 5 ints (reef_batch_plans) */
    int nplans;
} reef_server_config;

enum reef_request_state
{
    REEF_REQ_QUEUED = 0, REEF_REQ_LIVE = 1, REEF_REQ_DONE = 2, REEF_REQ_FAILED = 3
};

int reef_server_create(const reef_server_config* cfg, reef_server** out);
/* stops after the current round; requests not done end REEF_REQ_FAILED */
void reef_server_destroy(reef_server* s);
/* message of the last failed reef_server_* call of this thread */
const char* reef_server_last_error(void);
/* copies the grid and fields (the caller may free them on return); REEF_E_NOMEM if the
 * tank alone exceeds the budget */
int reef_server_submit(reef_server* s, const reef_grid_desc* grid, double dt, int steps, const double* phi,
                       const double* u, const double* v, const double* w, const double* press, long* id);
/* waits until the request is done or failed, or timeout_s passed (< 0: no limit);
 * state receives its reef_request_state; a failed request returns its error */
int reef_server_wait(reef_server* s, long id, double timeout_s, int* state);
/* a field (REEF_U .. REEF_VISC) of a done request after its last step */
int reef_server_result(reef_server* s, long id, int field_id, double* host);
/* times[3] = submitted, admitted, done; steps done; summed BiCGStab iterations */
int reef_server_request_info(reef_server* s, long id, double* times, int* steps_done, long* iters);
/* forgets a done or failed request */
int reef_server_release(reef_server* s, long id);
/* queued and live requests, scheduler rounds run, device bytes held by live tanks; returns the
 * worker's error once it has stopped on one */
int reef_server_stats(reef_server* s, int* queued, int* live, long* steps, double* bytes_in_use);
/* the largest device byte count the live tanks held at once since the server was created */
int reef_server_peak_bytes(reef_server* s, double* peak);

#ifdef __cplusplus
}
#endif

#endif
