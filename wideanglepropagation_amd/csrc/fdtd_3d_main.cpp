// fdtd_3d_bin: the reference program's command-line and file contract over libwapfdtd.so.
//
//   fdtd_3d_bin N_FRAMES STEPS_PER_FRAME SNAPSHOT_COUNT [-lambda] [PARAMS_FILE]
//
// Reads the potential from ./potential.bin (int32 nx, ny, nz + float32 V in
// (nz,ny,nx) order) and writes ./snapshots.bin (int32 nx, ny, nz, M, float32
// dt, E_sim, then M records of float32 t, phi[nz,ny,nx], psi[nz,ny,nx]),
// exactly as fdtd_3d.cu:284-306 and :346-394 do.  The grid and physics come
// from the #define lines of PARAMS_FILE (default ./fdtd_3d_params.h, else
// ./fdtd_3d.cu), the file RelativisticSolver3D._update_macros rewrites, read
// at run time instead of compiled in.  Constants are parsed as float, like
// the float literals of the reference (E_sim = LAP_COEFF*P_INF*P_INF in float,
// fdtd_3d.cu:316; t = float((frame*SPF + s) * DT), :365).  Unlike the
// reference, a short potential.bin is an error (fdtd_3d.cu:292-295 only warns).
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

int main(int argc, char **argv) {
  int n_frames = 200, spf = 50, m_snap = 10;
  if (argc > 1) n_frames = atoi(argv[1]);
  if (argc > 2) spf = atoi(argv[2]);
  if (argc > 3) m_snap = atoi(argv[3]);
  std::string params = "fdtd_3d_params.h";
  if (argc > 5) params = argv[5];
  else {
    std::ifstream t(params);
    if (!t.good()) params = "fdtd_3d.cu";
  }
  auto d = read_defines(params);
  if (d.empty()) {
    printf("Error: no #define parameters in %s.\n", params.c_str());
    return 1;
  }
  wap_desc desc;
  memset(&desc, 0, sizeof(desc));
  desc.nx = (int)num(d, "NX", 64);
  desc.ny = (int)num(d, "NY", 64);
  desc.nz = (int)num(d, "NZ", 203);
  desc.equation = WAP_EQ_SCHRODINGER;
  desc.precision = WAP_PREC_F32;
  desc.time_mode = WAP_TIME_F32;
  desc.src_enabled = 1;
  desc.src_k = (int)num(d, "K_SOURCE", 31);
  desc.src_band = (int)num(d, "SRC_BAND", 3);
  desc.src_alpha = num(d, "SRC_ALPHA", 0.05f);
  desc.xmin = num(d, "XMIN", -1.f);
  desc.xmax = num(d, "XMAX", 1.f);
  desc.ymin = num(d, "YMIN", -1.f);
  desc.ymax = num(d, "YMAX", 1.f);
  desc.zmin = num(d, "ZMIN", -1.6f);
  desc.zmax = num(d, "ZMAX", 1.5811f);
  const float dt = num(d, "DT", 0.0000005f);
  const float lap = num(d, "LAP_COEFF", 3.427719f);
  const float pinf = num(d, "P_INF", 125.663706f);
  const float e_sim = lap * pinf * pinf;
  const float k_beam = sqrtf(e_sim / lap);
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
    printf("Error: potential.bin not found.\n");
    return 1;
  }
  int hdr[3] = {0, 0, 0};
  size_t got_h = fread(hdr, sizeof(int), 3, fp);
  size_t got = fread(V.data(), sizeof(float), n, fp);
  fclose(fp);
  if (got_h != 3 || got != n || hdr[0] != desc.nx || hdr[1] != desc.ny || hdr[2] != desc.nz) {
    printf("Error: potential.bin holds %zu floats for a %dx%dx%d header; expected %zu (%dx%dx%d).\n",
           got, hdr[0], hdr[1], hdr[2], n, desc.nx, desc.ny, desc.nz);
    return 1;
  }

  printf("Initializing 3D Simulation (X, Y Periodic; Z PML)...\n");
  printf("Grid: %d x %d x %d\n", desc.nx, desc.ny, desc.nz);
  printf("E_SIM: %.4f (LapCoeff=%.4f)\n", e_sim, lap);

  wap_handle *h = nullptr;
  if (wap_create(&desc, &h)) return die("wap_create");
  if (wap_set_potential(h, V.data(), WAP_DT_F32, 0)) return die("wap_set_potential");

  FILE *fs = fopen("snapshots.bin", "wb");
  if (!fs) {
    printf("Error opening snapshots.bin\n");
    wap_destroy(h);
    return 1;
  }
  const int nx = desc.nx, ny = desc.ny, nz = desc.nz;
  fwrite(&nx, sizeof(int), 1, fs);
  fwrite(&ny, sizeof(int), 1, fs);
  fwrite(&nz, sizeof(int), 1, fs);
  fwrite(&m_snap, sizeof(int), 1, fs);
  fwrite(&dt, sizeof(float), 1, fs);
  fwrite(&e_sim, sizeof(float), 1, fs);

  std::vector<float> st(2 * n), phi(n), psi(n);
  int save_start = n_frames - m_snap;
  if (save_start < 0) save_start = 0;
  int saved = 0;
  for (int frame = 0; frame < n_frames; ++frame) {
    if (wap_step(h, spf, (int64_t)frame * spf, 0.0)) return die("wap_step");
    if (frame % 10 == 0) printf("Frame %d / %d\n", frame, n_frames);
    if (frame >= save_start && saved < m_snap) {
      if (wap_get_state(h, st.data(), nullptr, WAP_DT_C64, 0)) return die("wap_get_state");
      for (size_t t = 0; t < n; ++t) {
        phi[t] = st[2 * t];
        psi[t] = st[2 * t + 1];
      }
      const float t_snap = (float)(frame * spf) * dt;
      fwrite(&t_snap, sizeof(float), 1, fs);
      fwrite(phi.data(), sizeof(float), n, fs);
      fwrite(psi.data(), sizeof(float), n, fs);
      ++saved;
    }
  }
  fclose(fs);
  wap_destroy(h);
  printf("Done. Saved %d snapshots.\n", saved);
  return 0;
}
