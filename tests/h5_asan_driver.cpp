// Host-only driver of the HDF5 layer (csrc/h5io.cpp) for a build under AddressSanitizer: creates a
// projection-banked and a slice-banked virtual dataset in the directory given as argv[1], writes a volume,
// reads a box back, reads the layout and the theta dataset, and takes two error paths.  Prints one line per
// layout; tests/test_io_hdf5_asan.py compiles and runs it.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include "mosaic_h5.h"
static int check(int rc, const char* what) {
  if (rc) printf("%s rc=%d: %s\n", what, rc, mm_h5_last_error());
  return rc;
}
int main(int argc, char** argv) {
  if (argc < 2) return 2;
  const std::string dir = argv[1];
  const std::string pp = dir + "/proj.h5", ps = dir + "/slice.h5", pn = dir + "/none.h5";
  const char* paths[] = {pp.c_str(), ps.c_str()};
  long long shape[3] = {20, 12, 16};
  long long vch[3] = {8, 5, 16};
  for (int stype = 0; stype < 2; ++stype) {
    const char* p = paths[stype];
    if (check(mm_h5_create(p, shape, vch, stype, 2, nullptr), "create")) return 1;
    std::vector<float> a(20 * 12 * 16);
    for (size_t i = 0; i < a.size(); ++i) a[i] = (float)i;
    long long st[3] = {0, 0, 0};
    if (check(mm_h5_write(p, st, shape, a.data()), "write")) return 1;
    long long st2[3] = {3, 2, 1}, cnt2[3] = {11, 7, 9};
    std::vector<float> b(11 * 7 * 9, -1.f);
    if (check(mm_h5_read(p, st2, cnt2, b.data()), "read")) return 1;
    int bad = 0;
    for (int i = 0; i < 11; ++i)
      for (int j = 0; j < 7; ++j)
        for (int k = 0; k < 9; ++k)
          if (b[(i * 7 + j) * 9 + k] != a[((i + 3) * 12 + j + 2) * 16 + k + 1]) ++bad;
    mm_h5_info info;
    if (check(mm_h5_info_get(p, &info), "info")) return 1;
    std::vector<float> th(20);
    for (int i = 0; i < 20; ++i) th[i] = 0.5f * i;
    if (check(mm_h5_write_theta(p, th.data(), 20), "theta w")) return 1;
    long long n = 0;
    if (check(mm_h5_read_theta(p, nullptr, &n), "theta n")) return 1;
    std::vector<float> th2(n);
    if (check(mm_h5_read_theta(p, th2.data(), &n), "theta r")) return 1;
    // error paths: box outside, bad path
    long long cnt3[3] = {30, 1, 1};
    int e1 = mm_h5_read(p, st, cnt3, b.data());
    int e2 = mm_h5_info_get(pn.c_str(), &info);
    printf("stype %d: mismatches %d, virtual %d, shape %lld %lld %lld, theta n %lld equal %d, errors %d %d\n", stype,
           bad, info.is_virtual, info.shape[0], info.shape[1], info.shape[2], n,
           (int)(memcmp(th.data(), th2.data(), 80) == 0), e1, e2);
  }
  return 0;
}
