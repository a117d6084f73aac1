/* Boundary plot files of the fault data (host code, libtmk_h5.so).
 *
 * Replaces AMRLevelLinElast::writeCustomPlotFile and writeThisBdryLevel
 * (srcLinElast/AMRLevelLinElast.cpp:1628-1756): one HDF5 file per plot step, named
 * "<prefix>boundary.<step %06d>.2d.hdf5", holding the fault data of every level that has
 * fault boxes, in Chombo's AMR HDF5 layout (Chombo's HDF5Handle, HDF5HeaderData and the
 * LevelData write of CH_HDF5; Chombo is not part of the reference tree).
 *
 * File contents:
 *   /                      attributes max_level, num_levels, iteration, time, num_components,
 *                          component_<n> (the fault variables' names)
 *   /Chombo_global         attributes SpaceDim (2), testReal (0.0)
 *   /level_<l>             attributes ref_ratio, dx, dt, time, prob_domain (the fault face box of
 *                          the level's domain, bdryLo(domain, 1, 1));
 *                          datasets boxes (the level's fault boxes, lo_i lo_j hi_i hi_j),
 *                          Processors, data:offsets=0, data:datatype=0;
 *   /level_<l>/data_attributes   attributes comps, objectType ("FArrayBox"), ghost (4, 0),
 *                          outputGhost (the requested (1, 1) clipped to the data's ghost: (1, 0))
 * data:datatype=0 holds, box after box, every component of the box grown by outputGhost in
 * Fortran order (i fastest); data:offsets=0 holds where each box starts.
 */
#ifndef TETEMOKO_H5_H
#define TETEMOKO_H5_H

#include "tetemoko_hip.h" /* error codes */

#ifdef __cplusplus
extern "C" {
#endif

/* One level's fault data (AMRLevelLinElast's m_BNew on m_bdryGrids, :1713-1749). */
typedef struct tmk_bdry_level {
    int level;            /* level index: the group is "level_<level>"                          */
    int ref_ratio;        /* refinement ratio to the next finer level                           */
    double dx, dt, time;  /* the level's cell size, step and time                               */
    int domain_nx;        /* cells of the level's domain in x: prob_domain = (0,0)-(nx-1,0)      */
    int nbox;             /* fault boxes (level boxes with lo_j == 0); 0 writes no group         */
    const int* boxes;     /* nbox x {lo_i, lo_j, hi_i, hi_j}; lo_j = hi_j = 0 (the fault faces)  */
    const double* const* data; /* nbox pointers, each ncomp x (hi_i - lo_i + 1 + 2*ghost_x)
                                  values (component-major, 4 x-ghosts each side, no y-ghosts) */
} tmk_bdry_level;

/* Writes the boundary plot file at path: the root header (max_level, num_levels =
 * finest_level + 1, iteration = step, time, the ncomp component names) and one group per
 * level with nbox > 0, levels in the order given (AMRLevelLinElast.cpp:1650-1697,1713-1756).
 * Boxes are written in Chombo's layout order (by lo_i).  Returns 0, TMK_E_ARG for bad
 * arguments, or TMK_E_IO when the HDF5 library reports an error. */
int tmk_h5_write_bdry_plot(const char* path, int max_level, int finest_level, int step, double time,
                           int ncomp, const char* const* names, int nlev, const tmk_bdry_level* levels);

/* "<prefix>boundary.<step %06d>.2d.hdf5" into out (AMRLevelLinElast.cpp:1658-1661).
 * Returns the length, or -1 when it does not fit in cap bytes. */
int tmk_h5_bdry_plot_name(const char* prefix, int step, char* out, int cap);

#ifdef __cplusplus
}
#endif

#endif
