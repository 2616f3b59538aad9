// swd_io.cpp -- the surface-wave data file (surfdataSC.dat) read into the
// arrays the GPU path consumes: SurfTime::read_swd_data
// (src/SWD/io.cpp:29-238) with read_receiver (:6-27).
//
// Layout of the file: four header blocks (Rayleigh phase, Rayleigh group,
// Love phase, Love group), each a line "nmode k_0 .. k_{nmode-1}" followed by
// nmode lines of periods; then, for every source station, a line
// "# lat lon mode period wavetype veltype" followed by its receivers
// "lat lon velocity".  Positions are stored as colatitude and longitude in
// radians, distances in km (haversine, R = 6371), observed times as
// distance / velocity, all float as in the reference.  Host code only.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

namespace {

struct Table {
  int nmode = 0, nt = 0;
  std::vector<int> kmax;  // periods per mode
  std::vector<float> t;   // nt x nmode, column-major, -1 where a mode has fewer
};

thread_local std::string g_err;

// spherical.cpp:9-26 delsph: float variables, the trigonometry in double
// (dlat * 0.5 and M_PI * 0.5 promote), results stored to float as there
float delsph(float c1, float l1, float c2, float l2) {
  const float R = 6371.0f;
  const float dlat = c2 - c1, dlon = l2 - l1;
  const float la1 = (float)(M_PI * 0.5 - (double)c1), la2 = (float)(M_PI * 0.5 - (double)c2);
  const float a = (float)(std::sin((double)dlat * 0.5) * std::sin((double)dlat * 0.5) +
                          std::sin((double)dlon * 0.5) * std::sin((double)dlon * 0.5) *
                              std::cos((double)la1) * std::cos((double)la2));
  const float c = (float)(2.0 * std::atan2(std::sqrt((double)a), std::sqrt((double)(1 - a))));
  return R * c;
}

// io.cpp:57-87: "nmode k..." then nmode lines of periods, parsed with
// strtof.  Mode im's row holds its kmax[im] periods; the slots past them
// stay -1, the padding get_nt_row stops at (SurfaceWave.cpp:157-168).  (The
// reference reads nt numbers from every row, so strtof's 0 past the end of a
// shorter row overwrites the padding and the table gains T = 0 entries.)
bool read_table(std::ifstream &in, Table &T) {
  std::string line;
  if (!std::getline(in, line)) return false;
  const char *p = line.c_str();
  char *e = nullptr;
  T.nmode = (int)std::strtof(p, &e);
  p = e;
  if (T.nmode < 0 || T.nmode > 1000) return false;
  T.kmax.assign(T.nmode, 0);
  T.nt = 0;
  for (int i = 0; i < T.nmode; ++i) {
    T.kmax[i] = (int)std::strtof(p, &e);
    p = e;
    T.nt = std::max(T.nt, T.kmax[i]);
  }
  if (T.nt < 0 || T.nt > 100000) return false;
  T.t.assign((size_t)T.nt * T.nmode, -1.0f);
  for (int im = 0; im < T.nmode; ++im) {
    if (!std::getline(in, line)) return false;
    p = line.c_str();
    for (int it = 0; it < T.kmax[im]; ++it) {
      T.t[(size_t)im * T.nt + it] = std::strtof(p, &e);
      p = e;
    }
  }
  return true;
}

}  // namespace

struct jdsg_swd_data {
  Table tab[4];                 // Rc, Rg, Lc, Lg
  std::vector<int> pair_off;    // receivers of pair p: [pair_off[p], pair_off[p+1])
  std::vector<int> wave, mode, period;
  std::vector<float> src_colat, src_lon;
  std::vector<float> rcv_colat, rcv_lon, dist, obst;
};

extern "C" const char *jdsg_io_last_error(void) { return g_err.c_str(); }

// SurfTime::read_swd_data (io.cpp:29-238).  wave type codes: 0 Rc, 1 Rg,
// 2 Lc, 3 Lg (wavetype 2 / veltype 0 is Rc, 2 / 1 Rg, 1 / 0 Lc, else Lg);
// period is the 0-based index into the mode's periods.  Blank lines are
// skipped.
extern "C" int jdsg_read_swd_data(const char *path, jdsg_swd_data **out) {
  if (!path || !out) {
    g_err = "read_swd_data: null argument";
    return 1;
  }
  std::ifstream in(path);
  if (!in.is_open()) {
    g_err = std::string("cannot open file ") + path;
    return 1;
  }
  const float DEG2RAD = (float)(M_PI / 180.);
  auto *D = new jdsg_swd_data;
  for (int w = 0; w < 4; ++w)
    if (!read_table(in, D->tab[w])) {
      g_err = "read_swd_data: bad period header";
      delete D;
      return 1;
    }
  D->pair_off.push_back(0);
  std::string line;
  bool have = false;
  while (std::getline(in, line)) {
    if (line.find_first_not_of(" \t\r") == std::string::npos) continue;
    have = true;
    break;
  }
  while (have) {
    char dummy = 0;
    float lat = 0, lon = 0;
    int mode = 0, period = 0, wavetp = 0, veltp = 0;
    if (line[0] != '#' ||
        std::sscanf(line.c_str(), "%c%f%f%d%d%d%d", &dummy, &lat, &lon, &mode, &period, &wavetp,
                    &veltp) != 7) {
      g_err = "read_swd_data: expected '# lat lon mode period wavetype veltype', got: " + line;
      delete D;
      return 1;
    }
    // io.cpp:190-191: (90.0 - lat) in double times the float DEG2RAD
    const float c1 = (float)((90.0 - (double)lat) * (double)DEG2RAD), l1 = lon * DEG2RAD;
    const int code = (wavetp == 2 && veltp == 0)   ? 0
                     : (wavetp == 2 && veltp == 1) ? 1
                     : (wavetp == 1 && veltp == 0) ? 2
                                                   : 3;
    have = false;
    while (std::getline(in, line)) {  // io.cpp:6-27 read_receiver
      if (line.find_first_not_of(" \t\r") == std::string::npos) continue;
      if (line[0] == '#') {
        have = true;
        break;
      }
      float rlat = 0, rlon = 0, v = 0;
      if (std::sscanf(line.c_str(), "%f%f%f", &rlat, &rlon, &v) != 3) {
        g_err = "read_swd_data: expected 'lat lon velocity', got: " + line;
        delete D;
        return 1;
      }
      const float c2 = (float)((90.0 - (double)rlat) * (double)DEG2RAD), l2 = rlon * DEG2RAD;
      const float d = delsph(c1, l1, c2, l2);
      D->rcv_colat.push_back(c2);
      D->rcv_lon.push_back(l2);
      D->dist.push_back(d);
      D->obst.push_back(d / v);
    }
    D->wave.push_back(code);
    D->mode.push_back(mode);
    D->period.push_back(period - 1);
    D->src_colat.push_back(c1);
    D->src_lon.push_back(l1);
    D->pair_off.push_back((int)D->dist.size());
  }
  *out = D;
  return 0;
}

extern "C" void jdsg_swd_data_sizes(const jdsg_swd_data *D, int *nmode, int *nt,
                                    long long *npairs, long long *nobs) {
  for (int w = 0; w < 4; ++w) {
    if (nmode) nmode[w] = D->tab[w].nmode;
    if (nt) nt[w] = D->tab[w].nt;
  }
  if (npairs) *npairs = (long long)D->wave.size();
  if (nobs) *nobs = (long long)D->dist.size();
}

// the period table of wave type w (0 Rc .. 3 Lg): nt x nmode floats,
// column-major, -1 past a mode's count; kmax[nmode] the counts
extern "C" int jdsg_swd_data_table(const jdsg_swd_data *D, int w, float *t, int *kmax) {
  if (w < 0 || w > 3) {
    g_err = "swd_data_table: wave type must be 0..3";
    return 1;
  }
  const Table &T = D->tab[w];
  if (t) std::memcpy(t, T.t.data(), sizeof(float) * T.t.size());
  if (kmax) std::memcpy(kmax, T.kmax.data(), sizeof(int) * T.kmax.size());
  return 0;
}

// per pair: receiver offsets (npairs + 1), wave code, mode, 0-based period
// index, source colatitude / longitude (radians); per measurement: receiver
// colatitude / longitude, distance (km), observed time (s).  Any pointer may
// be null.
extern "C" int jdsg_swd_data_arrays(const jdsg_swd_data *D, int *pair_off, int *wave, int *mode,
                                    int *period, float *src_colat, float *src_lon,
                                    float *rcv_colat, float *rcv_lon, float *dist, float *obst) {
  auto cp = [](void *dst, const void *src, size_t bytes) {
    if (dst && bytes) std::memcpy(dst, src, bytes);
  };
  const size_t np = D->wave.size(), no = D->dist.size();
  cp(pair_off, D->pair_off.data(), sizeof(int) * (np + 1));
  cp(wave, D->wave.data(), sizeof(int) * np);
  cp(mode, D->mode.data(), sizeof(int) * np);
  cp(period, D->period.data(), sizeof(int) * np);
  cp(src_colat, D->src_colat.data(), sizeof(float) * np);
  cp(src_lon, D->src_lon.data(), sizeof(float) * np);
  cp(rcv_colat, D->rcv_colat.data(), sizeof(float) * no);
  cp(rcv_lon, D->rcv_lon.data(), sizeof(float) * no);
  cp(dist, D->dist.data(), sizeof(float) * no);
  cp(obst, D->obst.data(), sizeof(float) * no);
  return 0;
}

extern "C" void jdsg_swd_data_free(jdsg_swd_data *D) { delete D; }

// ---- the model file (MOD) ----
//
// DSurfTomo::read_model (src/surftomo/io.cpp:27-85; JSurfGTomo::read_model,
// src/JSurfGTomo/JSurfGtomo.cpp:57- reads the same layout): line 1
// "nx ny nz" (nlat nlon nz), line 2 "goxd gozd" (origin lat, lon), line 3
// "dvxd dvzd" (spacing in lat, lon), line 4 the nz depths (strtof), then the
// nx*ny*nz velocities read as whitespace-separated floats with i (lat)
// fastest, then j (lon), then k (depth): the column-major fmat3 vs(nx,ny,nz)
// that the GPU batch takes as it is.  Grid: lat[i] = goxd - i*dvxd,
// lon[j] = gozd + j*dvzd (io.cpp:84-89), both float.  Anything after a header line's
// numbers (the example's "# nlat nlon nz" comments) is ignored, as sscanf
// ignores it there.

namespace {

bool read_header(std::ifstream &in, const char *path, int *nx, int *ny, int *nz, float *g) {
  std::string l1, l2, l3;
  if (!std::getline(in, l1) || !std::getline(in, l2) || !std::getline(in, l3)) {
    g_err = std::string("read_model: header shorter than three lines in ") + path;
    return false;
  }
  if (std::sscanf(l1.c_str(), "%d%d%d", nx, ny, nz) != 3 ||
      std::sscanf(l2.c_str(), "%f%f", &g[0], &g[1]) != 2 ||
      std::sscanf(l3.c_str(), "%f%f", &g[2], &g[3]) != 2) {
    g_err = std::string("read_model: bad header in ") + path;
    return false;
  }
  if (*nx <= 0 || *ny <= 0 || *nz <= 0 || (long long)*nx * *ny * *nz > (1LL << 34)) {
    g_err = "read_model: grid dimensions must be positive";
    return false;
  }
  return true;
}

}  // namespace

extern "C" int jdsg_read_model_header(const char *path, int *nx, int *ny, int *nz, float *goxd,
                                      float *gozd, float *dvxd, float *dvzd) {
  if (!path || !nx || !ny || !nz || !goxd || !gozd || !dvxd || !dvzd) {
    g_err = "read_model_header: null argument";
    return 1;
  }
  std::ifstream in(path);
  if (!in.is_open()) {
    g_err = std::string("cannot open ") + path;
    return 1;
  }
  float g[4];
  if (!read_header(in, path, nx, ny, nz, g)) return 1;
  *goxd = g[0];
  *gozd = g[1];
  *dvxd = g[2];
  *dvzd = g[3];
  return 0;
}

extern "C" int jdsg_read_model(const char *path, int nx, int ny, int nz, float *dep, float *vs) {
  if (!path || !dep || !vs) {
    g_err = "read_model: null argument";
    return 1;
  }
  std::ifstream in(path);
  if (!in.is_open()) {
    g_err = std::string("cannot open ") + path;
    return 1;
  }
  int fx, fy, fz;
  float g[4];
  if (!read_header(in, path, &fx, &fy, &fz, g)) return 1;
  if (fx != nx || fy != ny || fz != nz) {
    g_err = "read_model: the file's grid is not the one the buffers were sized for";
    return 1;
  }
  // io.cpp:59-70: strtof over the depth line; a missing number reads as 0
  std::string line;
  if (!std::getline(in, line)) {
    g_err = "read_model: no depth line";
    return 1;
  }
  const char *p = line.c_str();
  char *e = nullptr;
  for (int k = 0; k < nz; ++k) {
    dep[k] = std::strtof(p, &e);
    p = e;
  }
  // io.cpp:73-78: i fastest; the reference leaves missing values
  // uninitialised, here a short file is an error
  const long long n = (long long)nx * ny * nz;
  for (long long q = 0; q < n; ++q) {
    if (!(in >> vs[q])) {
      g_err = "read_model: expected " + std::to_string(n) + " velocities, read " +
              std::to_string(q);
      return 1;
    }
  }
  return 0;
}

// The model snapshot the drivers write per iteration (results/mod_iter*.dat,
// src/surftomo/main.cpp:53-59 and :113-120, JSurfGTomo/main.cpp:66-71): one
// line "lon lat dep vs" per node, "%g" each, k slowest, then j, then i.
extern "C" int jdsg_write_model(const char *path, int nx, int ny, int nz, float goxd, float gozd,
                                float dvxd, float dvzd, const float *dep, const float *vs) {
  if (!path || !dep || !vs || nx <= 0 || ny <= 0 || nz <= 0) {
    g_err = "write_model: bad argument";
    return 1;
  }
  FILE *fp = std::fopen(path, "w");
  if (!fp) {
    g_err = std::string("cannot open ") + path + " for writing";
    return 1;
  }
  std::vector<float> lat(nx), lon(ny);
  for (int i = 0; i < nx; ++i) lat[i] = goxd - i * dvxd;
  for (int j = 0; j < ny; ++j) lon[j] = gozd + j * dvzd;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i)
        std::fprintf(fp, "%g %g %g %g\n", lon[j], lat[i], dep[k],
                     vs[i + (size_t)nx * (j + (size_t)ny * k)]);
  const bool ok = std::ferror(fp) == 0;
  if (std::fclose(fp) != 0 || !ok) {
    g_err = std::string("write_model: write failed on ") + path;
    return 1;
  }
  return 0;
}
