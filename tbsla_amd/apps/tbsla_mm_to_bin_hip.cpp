// tbsla_mm_to_bin_hip -- src/cpp/apps/tbsla_mm_to_binCOO.cpp on the GPU
// classes: reads a Matrix Market file into a COO block (readMM) and writes
// it in the binary format of MatrixCOO::write, then reads the file back and
// checks that the geometry survives the round trip.
//
//   tbsla_mm_to_bin_hip <file.mtx> <file.bin>
#include <tbsla/hip/Matrix.hpp>

#include <cstdio>
#include <fstream>

int main(int argc, char **argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: %s <file.mtx> <file.bin>\n", argv[0]);
    return 1;
  }
  tbsla::hip::MatrixCOO m;
  m.readMM(argv[1]);
  {
    std::ofstream os(argv[2], std::ofstream::binary);
    m.write(os);
  }
  std::printf("== read m2 ==\n");
  std::ifstream is(argv[2], std::ifstream::binary);
  tbsla::hip::MatrixCOO m2;
  m2.read(is);
  const tbsla_hip_geometry a = m.geometry(), b = m2.geometry();
  if (a.n_row != b.n_row || a.n_col != b.n_col || a.nnz != b.nnz) {
    std::fprintf(stderr, "round trip changed the geometry\n");
    return 2;
  }
  std::printf("n_row %d n_col %d nnz %lld\n", a.n_row, a.n_col, a.nnz);
  return 0;
}
