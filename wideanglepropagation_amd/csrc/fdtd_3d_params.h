/* Run parameters in the #define form of the reference's fdtd_3d.cu:17-67.
 * RelativisticSolver3D._update_macros rewrites the values of these lines the
 * way fdtd_solver.py:285-374 rewrites them in the CUDA source; the in-process
 * solver reads its parameters from Python, the fdtd_3d_bin command reads
 * this file's values from its --params argument. */
#define NX 64
#define NY 64
#define NZ 203

#define XMIN -1.0000f
#define XMAX  1.0000f
#define YMIN -1.0000f
#define YMAX  1.0000f
#define ZMIN -1.6000f
#define ZMAX  1.5811f

#define DT 0.0000005f

#define LAP_COEFF 3.427719f
#define P_INF 125.663706f

#define K_SOURCE 31
#define SRC_BAND 3
#define SRC_ALPHA 0.05f

#define PML_THICK   0.4000f
#define PML_THETA   0.7850f
#define PML_SIGMA0  80.0000f
#define PML_POWER   3.0f
