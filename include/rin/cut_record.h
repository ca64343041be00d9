// Per-tet cut record: the compact hand-off format of one tetrahedron's cell
// complex (the reference keeps a heap-allocated simplicial_arrangement
// Arrangement<3> / MaterialInterface<3> per active tet,
// src/implicit_arrangement.cpp:279-283, src/material_interface.cpp:323-327).
//
// A record is a sequence of uint16 words:
//   [0] kind          1 = implicit arrangement, 2 = material interface
//   [1] nlabels       4 + number of active functions (planes / materials)
//   [2] nv            vertices
//   [3] nf            faces
//   [4] nc            cells
//   [5] flags         bit 0: some plane/material duplicates another
//   [6] length        total words of the record, header included
//   [7] 0
//   vertices: nv x 4 words.  IA: three plane ids then 0xFFFF; MI: four labels.
//             Ids 0..3 are the tet faces (face i is opposite corner i),
//             4.. the active functions in ascending function order.
//   faces:    for each face: [a, b, c, n, v_0 .. v_{n-1}]
//             IA: a = supporting plane, b = positive cell, c = negative cell
//                 (0xFFFF on the tet boundary); the cycle's normal points to
//                 the positive cell.
//             MI: a = positive label, b = negative label, c = 0xFFFF; tet
//                 boundary faces have a in 0..3.
//   cells:    MI only: nc words, the material label of each cell.
//             IA cells are implied: cell c owns the faces naming it.
//   groups:   only when flags bit 0: nlabels words, the representative of
//             each label's coincidence group; IA then adds nlabels words of
//             orientation (1: same positive side as the representative).
#pragma once
#include <cstdint>

#define RIN_REC_NONE 0xFFFFu
#define RIN_REC_HEADER 8u
