/* Run parameters of the 2D solver in the #define form of the reference's
 * fdtd.cu (NX NY XMIN..YMAX PML_* J_SOURCE).  RelativisticSolver._update_macros
 * rewrites these lines the way fdtd_solver.py:54-99 does; fdtd_3d_bin reads
 * them at run time when the file has no NZ line. */
#define NX 1024
#define NY 1398

#define XMIN -1.0000f
#define XMAX  1.0000f
#define YMIN -2.0500f
#define YMAX  2.0468f

#define DT 0.000001f

#define J_SOURCE 156
#define SRC_BAND 3
#define SRC_ALPHA 0.05f

#define PML_THICK   0.4000f
#define PML_THETA   0.7850f
#define PML_SIGMA0  80.0000f
#define PML_POWER   3.0f
