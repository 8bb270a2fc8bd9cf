// Launchers of the fiber-tile kernels of the fused RK stage (fused_tile.hip).
#pragma once
#include "kern.hpp"

namespace adg {

// ---- fiber-tile kernels (fused_tile.hip) ----
// Fiber order of one dimension.  tile_prefix: pre[w] = set bits of the
// occupancy words before word w (nwords + 1 entries; cnt is scratch of the
// same size).  tile_order: perm[position] = element, and the tiles: the
// fibers whose first position lies in [t*te, (t+1)*te) form tile t, whose
// positions are [tbeg[t], tend[t]) (empty: tend < tbeg).
size_t tile_scan_bytes(int64_t nwords);
void tile_prefix(const uint32_t* bits, int64_t nwords, int32_t* cnt, int32_t* pre, void* scan_tmp,
                 size_t scan_tmp_bytes, hipStream_t s);
void tile_order(const uint32_t* bits, const int32_t* pre, const int32_t* fbase, const int32_t* slot, int64_t E, int ns,
                int te, int64_t ntile, int32_t* perm, int32_t* tbeg, int32_t* tend, hipStream_t s);
// Every task of the launch is evaluated on every tile of its swept dimension
// (ntile tiles per task); max_tile bounds the elements of a tile; tbeg / tend
// per dimension.
// allr: one thread per target accumulates four rest indices of a chunk (else
// one thread per (target, rest index)); last: the tasks are last sweeps
// (launched as their own kernel instance, so traces tell them apart).
template <class T>
void tile_sweep(int pf1, int pt1, const TaskDev* tasks, int ntask, int64_t ntile, int max_tile,
                const int32_t* const* tbeg, const int32_t* const* tend, int D, int cap, bool allr, bool last,
                hipStream_t s);
// Rest indices staged per chunk by tile_sweep for a term with R rest indices
// (0: the tile does not fit in LDS)
int tile_rest_chunk(int pf1, int R, int max_tile, int tbytes);
// up[e,q] = sum over the nterm term values (in order) and the point flux; the
// values of term j start j double-slots of n values into tv (for T = float as well)
template <class T>
void up_combine(const T* tv, int nterm, int64_t n, T* up, T* fp, const double* vcoef, int kind, double fs, int ncomp,
                hipStream_t s);
// rhs[e,q] = sum over the nterm term values (in order), then the RK stage
// update into u_next; non-finite stage values raise ERR_NONFINITE
template <class T>
void rhs_combine(const T* tv, const double* pv, int nterm, int npen, int64_t E, int K, double* rhs, const double* u,
                 const double* tn, double* u_next, int scheme, int stage, double dt, const double* dtv,
                 const int32_t* inst, int32_t* err, hipStream_t s);

}  // namespace adg
