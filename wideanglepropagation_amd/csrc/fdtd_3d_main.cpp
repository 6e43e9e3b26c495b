// fdtd_3d_bin: the reference programs' command-line and file contract over libwapfdtd.so.
//
//   fdtd_3d_bin N_FRAMES STEPS_PER_FRAME SNAPSHOT_COUNT [E_KINETIC | -lambda] [PARAMS_FILE]
//
// The grid and physics come from the #define lines of PARAMS_FILE (default
// ./fdtd_3d_params.h, ./fdtd_params.h, ./fdtd_3d.cu, ./fdtd.cu, the first that
// exists), the file RelativisticSolver3D / RelativisticSolver._update_macros
// rewrites, read at run time instead of compiled in.  Constants are parsed as
// float, like the float literals of the reference.  Unlike the reference, a
// short potential.bin is an error (fdtd_3d.cu:292-295, fdtd.cu:337-340 only warn).
//
// 3D (the file defines NZ), as fdtd_3d.cu:275-403: ./potential.bin = int32 nx,
// ny, nz + float32 V (nz,ny,nx); ./snapshots.bin = int32 nx, ny, nz, M, float32
// dt, E_sim, then M records of float32 t, phi[nz,ny,nx], psi[nz,ny,nx].
// E_sim = LAP_COEFF*P_INF*P_INF in float (fdtd_3d.cu:316), argv[4] is unused.
//
// 2D (no NZ), as fdtd.cu:323-552: ./potential.bin = int32 nx, ny + float32 V
// (ny,nx); ./snapshots.bin = int32 nx, ny, M, float32 dt, E_sim, then M records
// of float32 t, phi[ny,nx], psi[ny,nx].  Atomic units, H = -0.5 Lap + V, y is
// the propagation axis with end rows that zero the stage outputs; P_INF from
// argv[4] (fdtd.cu:356-380: < 0 is a wavelength in Angstrom, else the kinetic
// energy in Hartree), E_sim = 0.5 P_INF^2, k_beam = sqrtf(2 E_sim).  The 2D
// grid runs as a 3D one with ny = 1 (x periodic, z = the reference's y).
//
// Both: t = float((frame*SPF + s) * DT) (fdtd_3d.cu:365, fdtd.cu:438), the
// snapshot of a frame is stamped with the frame's start time.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/wap_fdtd.h"

static std::map<std::string, std::string> read_defines(const std::string &path) {
  std::map<std::string, std::string> out;
  std::ifstream f(path);
  std::string line;
  while (std::getline(f, line)) {
    std::istringstream is(line);
    std::string tok, name, val;
    if (!(is >> tok) || tok != "#define") continue;
    if (!(is >> name)) continue;
    if (!(is >> val)) continue;
    out[name] = val;
  }
  return out;
}

static float num(const std::map<std::string, std::string> &d, const char *k, float dflt) {
  auto it = d.find(k);
  if (it == d.end()) return dflt;
  std::string v = it->second;
  if (!v.empty() && (v.back() == 'f' || v.back() == 'F')) v.pop_back();
  return strtof(v.c_str(), nullptr);
}

static int die(const char *what) {
  fprintf(stderr, "Error: %s: %s\n", what, wap_last_error());
  return 1;
}

static bool exists(const std::string &p) {
  std::ifstream t(p);
  return t.good();
}

// P_INF of the 2D program from argv[4] (fdtd.cu:356-380), in float.
static float p_inf_2d(float e_kin_or_lambda) {
  const float c = 137.035999f, m = 1.0f;
  const float term_mass = m * c * c;
  if (e_kin_or_lambda < 0) {
    const float lambda_au = -e_kin_or_lambda / 0.5291772109f;
    return (2.0f * 3.14159265359f) / lambda_au;
  }
  const float e_total = term_mass + e_kin_or_lambda;
  return sqrtf(e_total * e_total - term_mass * term_mass) / c;
}

int main(int argc, char **argv) {
  int n_frames = 200, spf = 50, m_snap = 10;
  if (argc > 1) n_frames = atoi(argv[1]);
  if (argc > 2) spf = atoi(argv[2]);
  if (argc > 3) m_snap = atoi(argv[3]);
  std::string params;
  if (argc > 5) params = argv[5];
  else {
    for (const char *c : {"fdtd_3d_params.h", "fdtd_params.h", "fdtd_3d.cu", "fdtd.cu"})
      if (exists(c)) { params = c; break; }
  }
  auto d = read_defines(params);
  if (d.empty()) {
    printf("fdtd_3d_bin: no #define parameters in %s\n", params.empty() ? "(none found)" : params.c_str());
    return 1;
  }
  const bool two_d = d.find("NZ") == d.end();
  wap_desc desc;
  memset(&desc, 0, sizeof(desc));
  desc.equation = WAP_EQ_SCHRODINGER;
  desc.precision = WAP_PREC_F32;
  desc.time_mode = WAP_TIME_F32;
  desc.src_enabled = 1;
  desc.src_band = (int)num(d, "SRC_BAND", 3);
  desc.src_alpha = num(d, "SRC_ALPHA", 0.05f);
  desc.xmin = num(d, "XMIN", -1.f);
  desc.xmax = num(d, "XMAX", 1.f);
  float dt, lap, e_sim, k_beam;
  if (two_d) {
    desc.nx = (int)num(d, "NX", 1024);
    desc.ny = 1;
    desc.nz = (int)num(d, "NY", 1398);
    desc.src_k = (int)num(d, "J_SOURCE", 156);
    desc.zmin = num(d, "YMIN", -2.05f);
    desc.zmax = num(d, "YMAX", 2.0468f);
    desc.end_planes = WAP_ENDS_ZERO;
    dt = num(d, "DT", 0.000001f);
    lap = 0.5f;
    const float pinf = p_inf_2d(argc > 4 ? (float)atof(argv[4]) : num(d, "E_KINETIC", 300.f));
    e_sim = 0.5f * pinf * pinf;
    k_beam = sqrtf(2.0f * e_sim);
  } else {
    desc.nx = (int)num(d, "NX", 64);
    desc.ny = (int)num(d, "NY", 64);
    desc.nz = (int)num(d, "NZ", 203);
    desc.src_k = (int)num(d, "K_SOURCE", 31);
    desc.ymin = num(d, "YMIN", -1.f);
    desc.ymax = num(d, "YMAX", 1.f);
    desc.zmin = num(d, "ZMIN", -1.6f);
    desc.zmax = num(d, "ZMAX", 1.5811f);
    dt = num(d, "DT", 0.0000005f);
    lap = num(d, "LAP_COEFF", 3.427719f);
    const float pinf = num(d, "P_INF", 125.663706f);
    e_sim = lap * pinf * pinf;
    k_beam = sqrtf(e_sim / lap);
  }
  desc.dt = dt;
  desc.lap_coeff = lap;
  desc.e_sim = e_sim;
  desc.k_beam = k_beam;
  desc.pml_thick = num(d, "PML_THICK", 0.4f);
  desc.pml_sigma0 = num(d, "PML_SIGMA0", 80.f);
  desc.pml_theta = num(d, "PML_THETA", 0.785f);
  desc.pml_power = num(d, "PML_POWER", 3.f);

  const size_t n = (size_t)desc.nx * desc.ny * desc.nz;
  std::vector<float> V(n);
  FILE *fp = fopen("potential.bin", "rb");
  if (!fp) {
    printf("fdtd_3d_bin: cannot open potential.bin\n");
    return 1;
  }
  const int nhdr = two_d ? 2 : 3;
  int hdr[3] = {0, 0, 0};
  size_t got_h = fread(hdr, sizeof(int), nhdr, fp);
  size_t got = fread(V.data(), sizeof(float), n, fp);
  fclose(fp);
  const bool hdr_ok = two_d ? (hdr[0] == desc.nx && hdr[1] == desc.nz)
                            : (hdr[0] == desc.nx && hdr[1] == desc.ny && hdr[2] == desc.nz);
  if (got_h != (size_t)nhdr || got != n || !hdr_ok) {
    printf("fdtd_3d_bin: potential.bin has %zu values after its [%d %d %d] header, not %zu\n",
           got, hdr[0], hdr[1], two_d ? 0 : hdr[2], n);
    return 1;
  }

  if (two_d) {
    printf("2D grid, periodic in x, absorbing layers at both ends of the propagation axis\n");
    printf("nx %d, ny %d\n", desc.nx, desc.nz);
    printf("carrier %.4f, wavenumber %.4f\n", e_sim, k_beam);
  } else {
    printf("3D grid, periodic in x and y, absorbing layers at both z ends\n");
    printf("nx %d, ny %d, nz %d\n", desc.nx, desc.ny, desc.nz);
    printf("carrier %.4f, Laplacian coefficient %.4f\n", e_sim, lap);
  }

  wap_handle *h = nullptr;
  if (wap_create(&desc, &h)) return die("wap_create");
  if (wap_set_potential(h, V.data(), WAP_DT_F32, 0)) return die("wap_set_potential");

  FILE *fs = fopen("snapshots.bin", "wb");
  if (!fs) {
    printf("fdtd_3d_bin: cannot create snapshots.bin\n");
    wap_destroy(h);
    return 1;
  }
  const int nx = desc.nx, ny = desc.ny, nz = desc.nz;
  fwrite(&nx, sizeof(int), 1, fs);
  if (two_d) {
    fwrite(&nz, sizeof(int), 1, fs);   // the 2D file's ny is the propagation axis
  } else {
    fwrite(&ny, sizeof(int), 1, fs);
    fwrite(&nz, sizeof(int), 1, fs);
  }
  fwrite(&m_snap, sizeof(int), 1, fs);
  fwrite(&dt, sizeof(float), 1, fs);
  fwrite(&e_sim, sizeof(float), 1, fs);

  std::vector<float> st(2 * n), phi(n), psi(n);
  int save_start = n_frames - m_snap;
  if (save_start < 0) save_start = 0;
  int saved = 0;
  for (int frame = 0; frame < n_frames; ++frame) {
    if (wap_step(h, spf, (int64_t)frame * spf, 0.0)) return die("wap_step");
    if (frame % 10 == 0) printf("frame %d of %d\n", frame, n_frames);
    if (frame >= save_start && saved < m_snap) {
      if (wap_get_state(h, st.data(), nullptr, WAP_DT_C64, 0)) return die("wap_get_state");
      for (size_t t = 0; t < n; ++t) {
        phi[t] = st[2 * t];
        psi[t] = st[2 * t + 1];
      }
      const float t_snap = (float)(frame * spf) * dt;
      if (fwrite(&t_snap, sizeof(float), 1, fs) != 1 || fwrite(phi.data(), sizeof(float), n, fs) != n ||
          fwrite(psi.data(), sizeof(float), n, fs) != n) {
        fprintf(stderr, "fdtd_3d_bin: writing snapshots.bin failed at frame %d\n", frame);
        fclose(fs);
        wap_destroy(h);
        return 1;
      }
      ++saved;
    }
  }
  fclose(fs);
  wap_destroy(h);
  printf("%d snapshots written\n", saved);
  return 0;
}
