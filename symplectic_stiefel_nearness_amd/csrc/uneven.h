// Uneven batches: targets of mixed (n, m) solved by one launch per stage. The plan is built
// on the host (capi.hip: ssn_uneven_plan), copied by the caller to the start of the device
// workspace, and read there by the kernels.
//
// Plan layout (bytes, each part 256-aligned): UPlanHeader, UTarget[count], UItem[gram_items]
// (Gram work: (target, chunk), grouped by order), UItem[rec_items] (recovery work: (target,
// first row)). Within a group, work items are sorted by decreasing cost, so the persistent
// grids take the largest first.
#pragma once
#include <stddef.h>

namespace ssn {

struct UTarget {
    const double* T;  long long sT_r;          // target [2n][m], row stride (elements)
    double* X;        long long sX_r;          // output [2n][m]
    double* omega;    long long ldo;           // output [m][m]
    long long* iters; int* status; int* counters;
    double* partial;                           // [chunks][2][mp][mp] Gram partials
    double* Cr;       double* Ci;              // [m][m] each
    double* dual;                              // target_doubles(mp) solver block
    double* Pf;       double* Qf;              // [mp][mp] each
    double* frag;                              // [4 mp^2] recovery fragments
    int n, m, mp, chunks, chunk_rows, rec_rows;
    int eigh_small, eigh_vlds;                 // eigensolver variant (block list size, V in LDS)
};

struct UItem { int target; int index; };

struct UPlanHeader {
    int count, gram_items, rec_items, max_mp;
    int gram_begin[9];                         // Gram items of order mp = 16 k: [gram_begin[k-1], gram_begin[k])
    long long targets_off, gram_off, rec_off;  // byte offsets of the arrays from the plan start
    long long eigh_lds, dual_lds;              // dynamic LDS bytes of the eigensolver / dual launches
    long long plan_bytes, ws_bytes;            // plan size; workspace size (plan copy included)
};

}  // namespace ssn
