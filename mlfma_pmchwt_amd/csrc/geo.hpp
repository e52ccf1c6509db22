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

}  // namespace mlfma
