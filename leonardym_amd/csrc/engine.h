// engine.h -- the serving engine (internal): continuous batching of solve
// requests on one lattice.
//
// A LaneEngine runs the even-odd mixed-precision CG (the driver of
// source/inverters/PreconditionedBiCGStab.cpp:43-91 around the batched CG of
// batch.hip) over a batch whose lanes are independent requests: each lane has
// its own kappa (and so its own clover blocks and their inverses), its own
// tolerance and its own stopping point.  Between chunks of iterations the
// caller harvests the lanes that have stopped (their solutions go to host
// memory) and admits queued requests into free lanes, widening the batch when
// every lane is taken.  A lane's arithmetic and every sum over its sites do
// not depend on the batch width or on the other lanes, and a lane starts at a
// chunk boundary (where the reliable-update slots line up), so a request gives
// the same bits alone and in any mix.
#pragma once

#include <cstddef>
#include <vector>

#include "../../include/lym.h"

namespace lym_internal {

struct LaneEngine;

struct LaneRequest {
    const double* src;  // host, the reference's spinor layout
    double* dst;        // host, written when the lane is harvested
    double kappa, tol;  // solve D x = b to |b - D x| <= tol |b|
    void* tag;
    // device alternatives (a spinor of the engine's context): when set they
    // replace src / dst; eps > 0 replaces the tolerance by the absolute stopping
    // rule |r|^2 < eps on the even-odd system
    const lym_spinor* src_dev = nullptr;
    lym_spinor* dst_dev = nullptr;
    double eps = 0;
};

struct LaneResult {
    void* tag;
    int error;  // LYM_OK, or the code the request failed with
    lym_solve_info info;
    unsigned max_lanes;  // the most requests in the batch while this one ran (itself included)
};

// device bytes of an engine of up to wmax lanes on a lattice (what
// engine_create allocates, within a few percent)
size_t engine_bytes(const int dims[4], int rep, int storage, int prec, int wmax);
// ctx: the context (geometry, stream) the engine launches on; g: the lattice's
// links, whose clover field-strength blocks must already be built
int engine_create(lym_ctx* ctx, const lym_gauge* g, double csw, int prec, unsigned max_steps, int wmax,
                  LaneEngine** out);
void engine_destroy(LaneEngine* e);
int engine_room(const LaneEngine* e);  // requests that can still be admitted
int engine_busy(const LaneEngine* e);  // lanes holding a request
// Admits n requests (n <= engine_room).  A request that cannot start (e.g. a
// singular clover block at its kappa) is returned in `failed`.
int engine_admit(LaneEngine* e, int n, const LaneRequest* reqs, std::vector<LaneResult>* failed);
// Runs one chunk of iterations for every busy lane.
int engine_run_chunk(LaneEngine* e);
// Lanes that have stopped: their solutions are reconstructed and written to
// the requests' host buffers, and the lanes are freed.
int engine_harvest(LaneEngine* e, std::vector<LaneResult>* done);

// a context on an existing stream (the stream is not destroyed with it)
int ctx_create_on_stream(const int dims[4], int rep, int bc_t, void* stream, lym_ctx** out);

}  // namespace lym_internal
