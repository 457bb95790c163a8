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
    mpk::DevBuf buf_colptr, buf_rowind, buf_incptr, buf_incslot, buf_incpos, buf_T;
};
