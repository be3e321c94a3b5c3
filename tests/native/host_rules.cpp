// host_rules.cpp -- drives the host-only rules of the C ABI (configuration
// validation, kernel-shape choice, workgroup size) and the HDF5 field layout
// helpers under AddressSanitizer + UndefinedBehaviorSanitizer on the CPU
// (built and run by tests/test_sanitizers.py).  Exits 0 when every check
// holds; a sanitizer finding aborts the process.
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "config_rules.hpp"
#include "hdf5_layout.hpp"

static int failures = 0;
#define CHECK(c)                                                  \
  do {                                                            \
    if (!(c)) {                                                   \
      std::fprintf(stderr, "%s:%d: %s\n", __FILE__, __LINE__, #c); \
      ++failures;                                                 \
    }                                                             \
  } while (0)

static lbm_config base() {
  lbm_config c;
  std::memset(&c, 0, sizeof(c));
  c.lattice = LBM_D3Q19;
  c.collision = LBM_BGK;
  c.forcing_scheme = LBM_FS_GUO;
  c.force = LBM_FORCE_NONE;
  c.equilibrium = LBM_EQ_MA3;
  c.nx = c.ny = c.nz = 16;
  c.tau = 0.6;
  for (int d = 0; d < 3; ++d) c.wavelength[d] = 1.0;
  c.storage = LBM_F32;
  c.compute = LBM_F32;
  c.shifted = 1;
  c.residual_tol = 1e-10;
  c.force_kmax = c.removal_kmax = 2;
  return c;
}

int main() {
  std::string why;
  // every field of the configuration at ordinary, boundary and hostile values
  const int ints[] = {std::numeric_limits<int>::min(), -1, 0, 1, 2, 3, 4, 9, 10, 11, 64,
                      1024, 1 << 20, std::numeric_limits<int>::max()};
  const double dbls[] = {-1.0, 0.0, 1e-300, 0.5, 0.6, 1e300,
                         std::numeric_limits<double>::infinity(),
                         std::numeric_limits<double>::quiet_NaN()};
  int valid = 0, invalid = 0;
  for (int field = 0; field < 16; ++field)
    for (int v : ints) {
      lbm_config c = base();
      int* slots[16] = {&c.lattice, &c.collision, &c.forcing_scheme, &c.force, &c.equilibrium,
                        &c.nx, &c.ny, &c.nz, &c.storage, &c.shifted, &c.compute,
                        &c.alpha_mode, &c.kernel, &c.force_kmin, &c.force_kmax,
                        &c.removal_kmax};
      *slots[field] = v;
      if (c.lattice == LBM_D2Q9) c.nz = 1;
      const int rc = lbm::validate_config(&c, why);
      CHECK(rc == LBM_OK || (rc == LBM_E_INVALID && !why.empty()));
      if (rc == LBM_OK) {
        ++valid;
        const int V = lbm::pick_vector_width(&c, [](int) { return true; });
        CHECK(V == 0 || V == 2 || V == 4 || V == 8);
        const int nk = c.lattice == LBM_D2Q9 ? c.ny : c.nz;
        CHECK(V == 0 || nk % V == 0);
        CHECK(lbm::pick_vector_width(&c, [](int) { return false; }) == 0);
        const int b = lbm::pick_block(nk);
        CHECK(b >= 64 && b <= 256 && b % 64 == 0);
      } else {
        ++invalid;
      }
    }
  for (double v : dbls) {
    lbm_config c = base();
    c.tau = v;
    const int rc = lbm::validate_config(&c, why);
    CHECK(rc == ((v > 0.0 && v < 1e301) ? LBM_OK : LBM_E_INVALID));
    c = base();
    c.force = LBM_FORCE_SINUSOIDAL;
    c.wavelength[1] = v;
    (void)lbm::validate_config(&c, why);
  }
  // cells = nx ny nz must stay below 2^31
  {
    lbm_config c = base();
    c.nx = 2048;
    c.ny = 2048;
    c.nz = 511;
    CHECK(lbm::validate_config(&c, why) == LBM_OK);
    c.nz = 512;  // exactly 2^31 cells
    CHECK(lbm::validate_config(&c, why) == LBM_E_INVALID);
  }
  // the headline configuration takes the vector kernel with 16-byte lanes
  {
    lbm_config c = base();
    c.nx = c.ny = c.nz = 256;
    CHECK(lbm::pick_vector_width(&c, [](int) { return true; }) == 4);
  }
  for (int nk = 1; nk <= 4096; ++nk) {
    const int b = lbm::pick_block(nk);
    CHECK(b >= 64 && b <= 256);
  }
  // padded layout: round trip, zero padding, dimensions of gSD::pLength()
  const int shapes[][4] = {{2, 9, 12, 1}, {2, 1, 1, 1}, {2, 7, 2, 1}, {3, 4, 6, 8},
                           {3, 1, 1, 1}, {3, 5, 3, 7}, {3, 2, 2, 33}};
  for (const auto& s : shapes) {
    lbm::Grid g;
    lbm::set_extents(g, s[0], s[0] == 2 ? 9 : 19, s[1], s[2], s[3]);
    const int last = s[0] == 2 ? s[2] : s[3];
    CHECK(g.dims[g.D - 1] == (lbm::hsize_t)(2 * (last / 2 + 1)));
    std::vector<double> dense((size_t)g.N), back((size_t)g.N, -1.0), pad;
    for (size_t k = 0; k < dense.size(); ++k) dense[k] = 1.0 + (double)k;
    lbm::to_padded(g, dense.data(), pad);
    CHECK(pad.size() == g.padded);
    double extra = 0.0;
    for (size_t r = 0; r < pad.size() / g.dims[g.D - 1]; ++r)
      for (size_t k = (size_t)last; k < g.dims[g.D - 1]; ++k)
        extra += pad[r * g.dims[g.D - 1] + k] * pad[r * g.dims[g.D - 1] + k];
    CHECK(extra == 0.0);
    lbm::from_padded(g, pad, back.data());
    CHECK(back == dense);
  }
  std::printf("host rules: %d valid and %d rejected configurations, %d failures\n", valid,
              invalid, failures);
  return failures ? 1 : 0;
}
