// filt_io.hpp -- the file writers of filt_io.cpp, for the asynchronous
// snapshot writer (csrc/filt_snapshot.hip).
#pragma once

#include <stddef.h>

#include <vector>

#include "../../include/filtration_hip.h"

namespace filt {
namespace io {

// Host arrays of one rank's slab (localNx*Ny*Nz values each).
struct View {
    const double *P1, *S2, *x, *y, *z, *roS1_old, *roS2_old, *u2x, *u2y, *u2z;
    const int* media;
};

// Destinations of a save-file read; a null x, y or z is skipped.
struct Out {
    double *P1, *S2, *x, *y, *z, *roS1_old, *roS2_old;
    int* media;
};

// One rank's block [int j][P1][S2][x][y][z][roS1_old][roS2_old][media]
// (main.cpp:523-532); truncate creates the file first (rank 0's fopen "wb").
int write_save(const char* path, const View& v, size_t n, int j, bool truncate);

// restore (main.cpp:540-569): the blocks of ranks 0..rank in turn, keeping
// rank's.
int read_save(const char* path, const filt_params* p, int rank, int size, const Out& o, int* j);

// save_data_plots (main.cpp:319-347): headers on rank 0, then rank's lines.
int write_plots(const char* dir, const View& v, const filt_params* p, int localNx, int rank,
                int size, double t);

}  // namespace io
}  // namespace filt
