/*
 * Host-side tessellation for the hfv path (C-ABI).
 *
 * Replaces the external crate meshless_voronoi 0.6.1 (Cargo.lock:434-437) at its call sites
 * Voronoi::build / build_partial (src/space.rs:238-262, :574-587).  The tessellation is not on the
 * device path; its output feeds hfv_set_mesh (include/hfv.h).
 */
#ifndef HFV_HOST_H
#define HFV_HOST_H

#include <stdint.h>

#include "hfv.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hfv_mesh hfv_mesh;

/* Voronoi::build(generators, anchor = 0, box, dimensionality, periodic) (src/space.rs:574-580).
 * loc: [3n] xyz interleaved (1D: sorted along x).  n_threads <= 0: all OpenMP threads. */
int hfv_voronoi_build(int dim, const double* box, int periodic, int64_t n, const double* loc, int n_threads,
                      hfv_mesh** out);
/* Voronoi::build_partial(generators, mask, anchor = 0, box, dimensionality, periodic) (src/space.rs:236-245,
 * called by Space::volume_calculation with the active mask): only the cells of masked generators are
 * built; a face between a built and an unbuilt cell comes from the built one, faces between two unbuilt
 * cells are absent, every cell's CSR lists its faces, and unbuilt cells' volume, centroid and search
 * radius are 0.  mask [n] (NULL: every cell, the same as hfv_voronoi_build). */
int hfv_voronoi_build_partial(int dim, const double* box, int periodic, int64_t n, const double* loc,
                              const uint8_t* mask, int n_threads, hfv_mesh** out);
/* Pointers into the mesh's own arrays, valid until hfv_mesh_free. */
int hfv_mesh_view(const hfv_mesh* mesh, hfv_mesh_arrays* out);
void hfv_mesh_free(hfv_mesh* mesh);

#ifdef __cplusplus
}
#endif
#endif
