// build.h -- model set-up on the device (build.hip); internal.
#pragma once
#include <vector>

#include "common.h"

namespace pfem {

constexpr int BUILD_MAXINC = 32;  // elements sharing one node (8 on a structured hex grid)
constexpr int BUILD_MAXNB = 64;   // neighbours of one node, itself included (27 on a structured hex grid)
enum { BUILD_OK = 0, BUILD_TOO_DENSE = 1, BUILD_TOO_LARGE = 2 };

// Node adjacency as the SELL-64 pattern of the 3x3-BSR K and the element scatter slots.
struct DevicePattern {
    int64_t nslices = 0, nslots = 0;
    dbuf<int64_t> slice_ptr;  // nslices + 1
    dbuf<int32_t> rowlen;     // nn
    dbuf<int32_t> colind;     // nslots
    dbuf<int32_t> diagpos;    // nn
    dbuf<int32_t> slots;      // 64 per element: slot of block (a, b)
};

// BUILD_TOO_DENSE: a node with more than BUILD_MAXINC elements or BUILD_MAXNB
// neighbours; BUILD_TOO_LARGE: indices beyond int32.  The caller then builds on the host.
int build_pattern_device(int64_t nn, int64_t ne, const int64_t *d_conn, DevicePattern &P, hipStream_t s);

// 1 and dims = (nx, ny, nz) when conn is box_mesh's numbering of an nx x ny x nz grid
// (h_conn0: the 8 nodes of element 0 on the host), else 0.
int box_dims_from_conn(int64_t nn, int64_t ne, const int64_t *h_conn0, const int64_t *d_conn, int64_t dims[3],
                       hipStream_t s);

// Parity colours of element_colors (Assembly.jl:172-188) for a box: colour per element,
// the elements listed colour by colour in increasing order, colour offsets in cptr.
// Returns the number of colours.
int parity_colours_device(const int64_t dims[3], int32_t *d_colour, int64_t *d_order, std::vector<int64_t> &cptr,
                          hipStream_t s);

// Cp^-1 = I at every Gauss point (Mesh.jl:132-138).
void fill_identity6(int64_t ngp, double *d_cp, hipStream_t s);

// box_mesh (Mesh.jl:30-59) with the seeded interior jitter, written on the device.
void box_mesh_device(double lx, double ly, double lz, int64_t nx, int64_t ny, int64_t nz, double jitter,
                     uint64_t seed, double *d_nodes, int64_t *d_conn, hipStream_t s);

}  // namespace pfem
