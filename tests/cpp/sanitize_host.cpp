// sanitize_host.cpp -- the library's host-side code under AddressSanitizer
// and UBSan (built against tests/cpp/mock_capi.cpp, no GPU): binary images
// (io.hip and the C++ class's read/write) and the Matrix Market parser
// (mmio.hip), on valid and malformed input.
//
//   sanitize_host <tmpdir> <format> <image> [<format> <image> ...]
// Prints "sanitize_host ok" when every check held; any sanitizer report
// aborts the run.
#include <tbsla/hip/Matrix.hpp>

#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

namespace {

int failures = 0;
void expect(bool ok, const std::string &what) {
  if (!ok) {
    std::printf("FAILED: %s (%s)\n", what.c_str(), tbsla_hip_last_error());
    failures++;
  }
}

std::vector<char> slurp(const std::string &path) {
  std::ifstream f(path, std::ios::binary);
  return std::vector<char>((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
}

std::vector<char> image_of(tbsla_hip_matrix_t m) {
  size_t len = 0;
  if (tbsla_hip_write(m, nullptr, 0, &len, nullptr) != TBSLA_HIP_SUCCESS) return {};
  std::vector<char> out(len);
  size_t got = 0;
  if (tbsla_hip_write(m, out.data(), out.size(), &got, nullptr) != TBSLA_HIP_SUCCESS || got != len)
    return {};
  return out;
}

void check_image(int fmt, const std::string &path) {
  const std::vector<char> img = slurp(path);
  expect(!img.empty(), "read " + path);
  tbsla_hip_matrix_t m;
  tbsla_hip_create(&m, fmt);
  expect(tbsla_hip_read(m, img.data(), img.size(), nullptr) == TBSLA_HIP_SUCCESS, "parse " + path);
  const std::vector<char> once = image_of(m);
  expect(once.size() == img.size(), "written image has the read image's size: " + path);
  tbsla_hip_matrix_t m2;
  tbsla_hip_create(&m2, fmt);
  expect(tbsla_hip_read(m2, once.data(), once.size(), nullptr) == TBSLA_HIP_SUCCESS, "re-parse");
  expect(image_of(m2) == once, "write(read(write(x))) == write(x): " + path);
  // the C++ class over the same calls, through streams
  {
    tbsla::hip::Matrix M(fmt);
    std::istringstream is(std::string(img.begin(), img.end()));
    M.read(is);
    std::ostringstream os;
    M.write(os);
    const std::string s = os.str();
    expect(std::vector<char>(s.begin(), s.end()) == once, "Matrix::write after Matrix::read");
  }
  // every strict prefix is refused (and read nothing past its end)
  const size_t step = std::max<size_t>(1, img.size() / 97);
  for (size_t cut = 0; cut < img.size(); cut += step) {
    std::vector<char> pre(img.begin(), img.begin() + (long)cut);  // exact-size buffer
    tbsla_hip_matrix_t t;
    tbsla_hip_create(&t, fmt);
    expect(tbsla_hip_read(t, pre.data(), pre.size(), nullptr) != TBSLA_HIP_SUCCESS,
           "truncated image refused at " + std::to_string(cut));
    tbsla_hip_destroy(t);
  }
  // a first array whose count is far beyond the image is refused
  {
    std::vector<char> bad = img;
    const size_t at = 6 * 4 + 8 + 4 * 4 + (fmt == TBSLA_HIP_ELL ? 4 : 0);
    const size_t huge = (size_t)1 << 60;
    std::memcpy(bad.data() + at, &huge, sizeof huge);
    tbsla_hip_matrix_t t;
    tbsla_hip_create(&t, fmt);
    expect(tbsla_hip_read(t, bad.data(), bad.size(), nullptr) != TBSLA_HIP_SUCCESS,
           "oversized array count refused");
    tbsla_hip_destroy(t);
  }
  tbsla_hip_destroy(m);
  tbsla_hip_destroy(m2);
}

void write_file(const std::string &path, const std::string &text) {
  std::ofstream f(path, std::ios::binary);
  f << text;
}

// parse a Matrix Market file; returns the status and the entry count
int mm(const std::string &path, long long *nnz) {
  int nr = 0, nc = 0;
  *nnz = -1;
  int st = tbsla_hip_mm_parse(path.c_str(), &nr, &nc, nnz, nullptr, nullptr, nullptr, 0);
  if (st != TBSLA_HIP_SUCCESS) return st;
  std::vector<int> r((size_t)*nnz), c((size_t)*nnz);
  std::vector<double> v((size_t)*nnz);
  long long n2 = 0;
  st = tbsla_hip_mm_parse(path.c_str(), &nr, &nc, &n2, r.data(), c.data(), v.data(), *nnz);
  for (long long k = 0; st == TBSLA_HIP_SUCCESS && k < n2; k++)
    if (r[k] < 0 || r[k] >= nr || c[k] < 0 || c[k] >= nc) st = -1;
  return st;
}

void check_matrix_market(const std::string &dir) {
  struct Case {
    const char *name, *text;
    long long nnz;  // expected entries, or -1: refused
  };
  const Case cases[] = {
      {"general.mtx",
       "%%MatrixMarket matrix coordinate real general\n% a comment\n3 4 3\n1 1 1.5\n2 3 -2e3\n3 4 7\n",
       3},
      {"crlf.mtx", "%%MatrixMarket matrix coordinate real general\r\n2 2 2\r\n1 1 1\r\n2 2 2\r\n", 2},
      {"blank.mtx", "%%MatrixMarket matrix coordinate real general\n\n2 2 1\n\n2 1 0.25\n\n", 1},
      {"symmetric.mtx",
       "%%MatrixMarket matrix coordinate real symmetric\n3 3 3\n1 1 4\n2 1 -1\n3 2 -1\n", 5},
      {"integer.mtx", "%%MatrixMarket matrix coordinate integer general\n2 2 2\n1 2 3\n2 1 4\n", -1},
      {"no_newline.mtx", "%%MatrixMarket matrix coordinate real general\n2 2 1\n2 2 9", 1},
      {"short.mtx", "%%MatrixMarket matrix coordinate real general\n3 3 4\n1 1 1\n2 2 2\n", -1},
      {"range.mtx", "%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1\n", -1},
      {"zero_index.mtx", "%%MatrixMarket matrix coordinate real general\n2 2 1\n0 1 1\n", -1},
      {"array.mtx", "%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n", -1},
      {"no_header.mtx", "2 2 1\n1 1 1\n", -1},
      {"empty.mtx", "", -1},
      {"garbage.mtx", "%%MatrixMarket matrix coordinate real general\n2 2 1\n1 x 1\n", -1},
  };
  for (const Case &c : cases) {
    const std::string path = dir + "/" + c.name;
    write_file(path, c.text);
    long long n = 0;
    const int st = mm(path, &n);
    if (c.nnz < 0) expect(st != TBSLA_HIP_SUCCESS, std::string("refused: ") + c.name);
    else expect(st == TBSLA_HIP_SUCCESS && n == c.nnz, std::string("parsed: ") + c.name);
  }
  long long n = 0;
  expect(mm(dir + "/does_not_exist.mtx", &n) != TBSLA_HIP_SUCCESS, "missing file refused");
  // a larger file, cut into the parser's threaded chunks
  std::string big = "%%MatrixMarket matrix coordinate real general\n5000 5000 200000\n";
  for (long long k = 0; k < 200000; k++)
    big += std::to_string(1 + (k * 7919) % 5000) + " " + std::to_string(1 + (k * 104729) % 5000) +
           " " + std::to_string(0.001 * k) + "\n";
  write_file(dir + "/big.mtx", big);
  expect(mm(dir + "/big.mtx", &n) == TBSLA_HIP_SUCCESS && n == 200000, "parsed: big.mtx");
}

int format_of(const std::string &f) {
  if (f == "COO") return TBSLA_HIP_COO;
  if (f == "SCOO") return TBSLA_HIP_SCOO;
  if (f == "CSR") return TBSLA_HIP_CSR;
  if (f == "ELL") return TBSLA_HIP_ELL;
  return TBSLA_HIP_DENSE;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 2) {
    std::printf("usage: sanitize_host <tmpdir> [<format> <image>]...\n");
    return 2;
  }
  for (int i = 2; i + 1 < argc; i += 2) check_image(format_of(argv[i]), argv[i + 1]);
  check_matrix_market(argv[1]);
  if (failures) {
    std::printf("%d checks failed\n", failures);
    return 1;
  }
  std::printf("sanitize_host ok\n");
  return 0;
}
