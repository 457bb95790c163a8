// CSC assembly plan: the structural sparse() pattern of a mesh + DOF map and,
// per column, the incident (element, local edge) pairs in ascending element
// order with the in-column position of each of the element's 6 rows.
#pragma once

#include "mpk_common.hpp"

// Columns handled by one wave64 block of the column-owner kernels.
constexpr int kTileCols = 64;

struct mpk_plan {
    int64_t ne = 0;            // elements
    int64_t nn = 0;            // nodes
    int64_t n = 0;             // columns == rows == NumGlobalDofs
    int64_t nnz = 0;           // structural nonzeros
    int64_t ninc = 0;          // incidences == diagonal triplets
    int64_t max_tile_nnz = 0;  // largest nnz over kTileCols consecutive columns
    int device = 0;
    int32_t* colptr = nullptr;    // [n+1]
    int32_t* rowind = nullptr;    // [nnz], 0-based
    int32_t* inc_ptr = nullptr;   // [n+1]
    uint32_t* inc_slot = nullptr; // [ninc], 6*e + c
    uint2* inc_pos = nullptr;     // [ninc], byte r = in-column position of row r (0xFF: none)
    int4* T = nullptr;            // [ne], 0-based node ids
    int32_t* cell = nullptr;      // [6*ne], 0-based DOF of local edge c (-1: inactive)
    mpk::DevBuf buf_colptr, buf_rowind, buf_incptr, buf_incslot, buf_incpos, buf_T, buf_cell;
    // Element state of the Newton kernels, [ne] x 48 B, allocated on first use:
    // double4 (Bx, By, Bz, nu) at state, then double2 (dnu/d|B|^2, sigma) at
    // state + 4*ne.
    mutable mpk::DevBuf buf_state;
    // Tile classes of the fused Newton set, written with the element state,
    // [2 ntiles + 1] int32 (ntiles = column tiles of kTileCols columns):
    // flag per tile (the call's epoch if an element with an M or D term
    // touches the tile), the list of flagged tiles, their count.  Allocated
    // (and cleared) on first use; each call takes the next epoch, so the
    // flags of earlier calls need no clearing.
    mutable mpk::DevBuf buf_tile_flag;
    mutable int32_t tile_epoch = 0;
    mutable int num_cus = 0;  // compute units of the plan's device (queried on first use)
    // side stream of the fused Newton set (its two tile classes run side by side)
    mutable mpk::SideStream side;

    // Closed-form (frame) kernels: one 16-byte record per incidence.
    //   x = element id
    //   y = c | orientation bits f_r' (bits 8..13) | pos'_0 << 16 | pos'_1 << 24
    //   z = pos'_2 | pos'_3 << 8 | pos'_4 << 16 | pos'_5 << 24
    //   w = 0
    // The frame orders the element's nodes (lo, hi, o1, o2): lo/hi are the
    // column edge's endpoints by global id (the same two nodes for every
    // incidence of the column), o1/o2 the other two in local order.  r' runs
    // over the frame's local edges, pos'_r' is the in-column position of that
    // edge's row (0xff: inactive) and f_r' = 1 when the frame edge runs from
    // the lower to the higher global node id.
    // Records are stored k-major per tile: the k-th incidence of the tile's
    // column `lane` sits at frame_base[tile] + 64 k + lane, so a wave reads the
    // k-th records of its 64 columns as one contiguous block.
    uint4* inc_frame = nullptr;
    int32_t* frame_base = nullptr;  // [ntiles+1]
    mpk::DevBuf buf_incframe, buf_framebase;

    // Ring records (closed-form modes, meshes with < 2^24 elements and nodes):
    // the incidences of each column in rotation order around the column edge,
    // so consecutive elements share a face and each record brings one new
    // node.  Same k-major tile layout and frame_base as inc_frame.  Frame of
    // record k: (lo, hi, o_prev, o_new), o_prev = o_new of record k-1 (record
    // 0: o_start).  Packing:
    //   x = element (24 bits) | f bits << 24      y = o_new (24 bits) | pos'_0 << 24
    //   z = pos'_1..pos'_4 (bytes)                w = pos'_5 | o_start << 8
    bool ring = false;  // every column is a simple ring or chain
    uint4* inc_ring = nullptr;
    int2* col_ends = nullptr;  // [n] (lo, hi) global node ids of each column edge
    mpk::DevBuf buf_incring, buf_colends;

    // Compact ring records (8 B per incidence), built from the ring records
    // when every field fits; the 16-byte ring records are then released.
    // Consecutive elements of a ring share the face (lo, hi, o_prev), so the
    // positions of the frame edges (lo, o_prev) and (hi, o_prev) are the
    // previous record's (lo, o_new) and (hi, o_new); the orientation bits
    // follow from the four node ids.  Per incidence (uint2, same k-major
    // tile layout and frame_base):
    //   x = element (24 bits) | pos'_2 << 24 (6 bits) | (pos'_5 & 3) << 30
    //   y = (o_new - lo) as 22-bit two's complement | pos'_4 << 22 (6 bits)
    //       | (pos'_5 >> 2) << 28
    // (6-bit positions, 63 = inactive row).  Per column (uint4):
    //   x = lo | pos'_0 << 24, y = hi | pos'_1 << 24, z = o_start | pos'_3 << 24
    //   (first record's pos'_1 / pos'_3, 8 bits, 0xff = inactive),
    //   w = colptr[j] - colptr[tile start] (16 bits) | incidences << 16 (8 bits).
    // Needs nodes < 2^24, |o_new - lo| < 2^21, column nnz <= 63, <= 255
    // incidences per column and tiles of <= 65535 entries.
    bool ring8 = false;
    uint2* inc_ring8 = nullptr;
    uint4* col_head = nullptr;
    mpk::DevBuf buf_incring8, buf_colhead;
    // Node coordinates padded to 32 bytes (x, y, z, 0), refreshed from the
    // caller's P by mpk_plan_set_coords and by every frame-kernel call given P.
    double4* P4 = nullptr;
    mutable bool coords_loaded = false;  // P4 holds the coordinates of some P
    mpk::DevBuf buf_P4;
};
