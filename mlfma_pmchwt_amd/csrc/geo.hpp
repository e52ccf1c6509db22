// Plain-old-data records shared by host setup code and device kernels.
#pragma once

namespace mlfma {

// Geometry of one RWG function, in the loop order of the reference's near-field
// operators: side 0 = negative face, side 1 = positive face
// (Code/Operator.h:727-776).
struct RwgGeo {
    double tri[2][9];    // v1, v2, v3 (xyz) of each face, file vertex order
    double vf[2][3];     // free vertex of each face
    double nrm[2][3];    // unit normal (snapped, Code/NASReader.cpp:52-56)
    double area[2];
    double len;          // common edge length
    int face[2];         // face ids, for the identical-triangle test
};

struct GaussRule {       // Dunavant rule, weights sum to 0.5
    int n;
    double l1[16], l2[16], l3[16], w[16];
};

// Near field of one target leaf t: `planes` dense column-major matrices of
// nt rows and C_t = sum of its source leaves' sizes columns, one after the
// other from tgt_off[t]; the columns are the source leaves in pair order (near
// neighbours, then t itself) and pair p starts at column pair_col[p].
// PEC: 1 plane; PMCHWT: z11, zK, z22 (z12 = -zK, z21 = zK).
struct NearDev {
    int planes, nleaf, npairs, N;
    const int* leaf_ptr;
    const int* perm;
    const int* order;         // target leaves by decreasing nt * C_t (launch order)
    const long long* tgt_off; // first entry of each target's plane 0
    const int* tgt_cols;      // C_t
    const int* col_ptr;       // per target, first entry in colref
    const int* colref;        // caller-order index of x for every column
    const int* pair_tgt;      // set-up only
    const int* pair_src;
    const int* pair_col;      // first column of the pair inside its target
    int nitem;                // matrix-core work items: (target, first 16-row tile), <= 4 tiles each,
    const int2* items;        // by decreasing rows * C_t
};

}  // namespace mlfma
