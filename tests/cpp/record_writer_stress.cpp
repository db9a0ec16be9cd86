// The generator's record writer (coulomb-explosion_amd/csrc/record_writer.hpp)
// under ThreadSanitizer, driven by the generator's own protocol: batch b is
// filled into buffer b % 2, submitted once filled, and buffer b % 2 is refilled
// only after wait_written(b - 1) says batch b - 2 has left it.  A buffer touched
// while the writer still reads it is a data race TSan reports; the file must
// hold every batch in order.  Also: a writer destroyed without finish(), and
// an output path that cannot be opened.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iterator>
#include <stdexcept>
#include <string>
#include <vector>

#include "record_writer.hpp"

namespace {

std::uint8_t byte_of(std::size_t batch, std::size_t i) {
  return static_cast<std::uint8_t>((batch * 131 + i * 7 + (i >> 8)) & 0xff);
}

int run(const std::string& path, std::size_t batches, std::size_t bytes) {
  std::vector<std::uint8_t> buf[2] = {std::vector<std::uint8_t>(bytes),
                                      std::vector<std::uint8_t>(bytes)};
  std::vector<std::size_t> size(batches);
  {
    ce::RecordWriter w(path);
    for (std::size_t b = 0; b < batches; ++b) {
      if (b >= 2) w.wait_written(b - 1);  // batch b - 2 has left buffer b % 2
      size[b] = bytes - (b * 977) % (bytes / 2);  // ragged batch sizes
      for (std::size_t i = 0; i < size[b]; ++i) buf[b % 2][i] = byte_of(b, i);
      w.submit(buf[b % 2].data(), size[b]);
    }
    if (w.submitted() != batches) return 1;
    w.finish();
  }
  std::ifstream in(path, std::ios::binary);
  const std::vector<std::uint8_t> got((std::istreambuf_iterator<char>(in)),
                                      std::istreambuf_iterator<char>());
  std::size_t off = 0;
  for (std::size_t b = 0; b < batches; ++b) {
    for (std::size_t i = 0; i < size[b]; ++i, ++off) {
      if (off >= got.size() || got[off] != byte_of(b, i)) {
        std::fprintf(stderr, "batch %zu byte %zu wrong\n", b, i);
        return 1;
      }
    }
  }
  return off == got.size() ? 0 : 1;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: record_writer_stress DIR\n");
    return 2;
  }
  const std::string dir = argv[1];
  if (run(dir + "/records.bin", 200, 1 << 16) != 0) return 1;
  if (run(dir + "/one.bin", 1, 4096) != 0) return 1;
  {  // destroyed without finish(): the thread stops after the queued buffer
    static std::uint8_t one[64] = {1};
    ce::RecordWriter w(dir + "/unfinished.bin");
    w.submit(one, sizeof(one));
  }
  try {
    ce::RecordWriter w(dir + "/no/such/dir/x.bin");
    std::fprintf(stderr, "opening a bad path did not throw\n");
    return 1;
  } catch (const std::runtime_error&) {
  }
  std::printf("record writer ok\n");
  return 0;
}
