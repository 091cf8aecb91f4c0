// fly::io::BinaryFile of the C++ facade (gsd.hpp:118-327, gsd.cpp:63-253): the reference's own round trip
// (tests/IO/gsd.cpp:32-111: box, TypeMap<Index>, typeid and positions over two frames, the second frame
// falling back to the first for chunks it lacks), then what that test leaves open: the file header at its
// fixed offsets, index and name-list growth past their first allocations, appending after a reopen,
// clear(), per-species values, a triclinic box, and the error paths (missing chunk, wrong rows, columns
// or type, writing a read-only file, a file that is not GSD, a frame never committed).
// Usage: gsd_check <directory>; leaves <directory>/layout.gsd for a reader written from the published
// format to check.  Prints one line per case; exit status 0 on success.  gsd_check --read <file> prints
// what it reads from a file another writer made.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "libfly/gpu/gsd.hpp"

using namespace fly;

static int failures = 0;
#define REQUIRE(c)                                                           \
  do {                                                                       \
    if (!(c)) {                                                              \
      std::fprintf(stderr, "REQUIRE failed: %s (line %d)\n", #c, __LINE__); \
      ++failures;                                                            \
    }                                                                        \
  } while (0)

template <typename F>
static bool throws_with(F&& f, char const* text) {
  try {
    f();
  } catch (RuntimeError const& e) {
    if (std::strstr(e.what(), text)) return true;
    std::fprintf(stderr, "unexpected message: %s\n", e.what());
    return false;
  }
  return false;
}

static std::vector<unsigned char> slurp(std::string const& path) {
  std::ifstream f(path, std::ios::binary);
  return {std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>()};
}

template <typename T>
static T at(std::vector<unsigned char> const& b, std::size_t off) {
  T v;
  std::memcpy(&v, b.data() + off, sizeof(T));
  return v;
}

using Cell = system::SoA<Position, TypeID, Frozen>;

// The reference's test case (tests/IO/gsd.cpp:32-111), with frame 1's positions checked against their
// values (the reference compares the read array with itself there).
static void reference_round_trip(std::string const& path) {
  Mat I = Mat::Zero();
  for (int k = 0; k < 3; ++k) I(k, k) = 1;
  system::Box box(I, Arr<bool>::Constant(true));
  system::TypeMap<Index> map(2);
  map.set(0, "Fe", 6);
  map.set(1, "H", 1);
  Cell cell(4);
  for (int i = 0; i < 4; ++i) {
    cell(r_, i) = Vec::Constant(i);
    cell(id_, i) = i == 0 ? 0u : 1u;
    cell(fzn_, i) = i == 2;
  }
  std::vector<double> initial(cell[r_].data(), cell[r_].data() + 12);
  {
    io::BinaryFile file(path, io::create);
    file.commit([&] {
      file.write(box);
      file.write(map);
      file.write("particles/N", static_cast<std::uint32_t>(cell.size()));
      file.write(id_, cell);
      file.write(r_, cell);
    });
    for (int i = 0; i < 4; ++i) cell(r_, i) += Vec::Constant(1);
    file.commit([&] {
      file.write("particles/N", static_cast<std::uint32_t>(cell.size()));
      file.write(r_, cell);
    });
  }
  io::BinaryFile file(path, io::read_only);
  REQUIRE(file.n_frames() == 2);
  auto out_map = file.read_map<Index>(0);
  REQUIRE(out_map.num_types() == 2);
  for (int i = 0; i < 2; ++i) {
    REQUIRE(out_map.get(i, Type{}) == map.get(i, Type{}));
    REQUIRE(out_map.get(i, Index{}) == map.get(i, Index{}));
  }
  system::Box out_box = file.read_box(0);
  double diff = 0;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) diff += std::pow(out_box.basis()(r, c) - box.basis()(r, c), 2);
  REQUIRE(std::sqrt(diff) < 1e-4);
  for (int k = 0; k < 3; ++k) REQUIRE(out_box.periodic(k));
  Cell out(file.read<std::uint32_t>(0, "particles/N"));
  REQUIRE(out.size() == 4);
  file.read_to(0, id_, out);
  file.read_to(0, r_, out);
  for (int i = 0; i < 4; ++i) REQUIRE(out(id_, i) == cell(id_, i));
  for (int k = 0; k < 12; ++k) REQUIRE(out[r_][static_cast<std::size_t>(k)] == initial[static_cast<std::size_t>(k)]);
  file.read_to(1, r_, out);
  for (int k = 0; k < 12; ++k) REQUIRE(out[r_][static_cast<std::size_t>(k)] == initial[static_cast<std::size_t>(k)] + 1);
  // Frame 1 holds no typeid chunk: it reads as frame 0's; so does a frame past the end.
  Cell out1(4);
  file.read_to(1, id_, out1);
  for (int i = 0; i < 4; ++i) REQUIRE(out1(id_, i) == cell(id_, i));
  REQUIRE(file.read<std::uint32_t>(7, "particles/N") == 4);
  std::printf("reference round trip ok (2 frames, fallback to frame 0)\n");
}

// Header fields at their offsets in the file, as the GSD 2.0 layout fixes them.
static void header_layout(std::string const& path) {
  auto b = slurp(path);
  REQUIRE(b.size() > 256 + 128 * 32 + 1024);
  REQUIRE(at<std::uint64_t>(b, 0) == 0x65DF65DF65DF65DFull);
  REQUIRE(at<std::uint64_t>(b, 8) == 256);                       // index location
  REQUIRE(at<std::uint64_t>(b, 16) == 128);                      // index entries allocated
  REQUIRE(at<std::uint64_t>(b, 24) == 256 + 128 * 32);           // name list location
  REQUIRE(at<std::uint64_t>(b, 32) == 1024);                     // name list bytes
  REQUIRE(at<std::uint32_t>(b, 40) == ((1u << 16) | 4u));        // schema version 1.4
  REQUIRE(at<std::uint32_t>(b, 44) == (2u << 16));               // file version 2.0
  REQUIRE(std::string(reinterpret_cast<char const*>(b.data() + 48)) == "openFLY");
  REQUIRE(std::string(reinterpret_cast<char const*>(b.data() + 112)) == "hoomd");
  std::printf("header layout ok\n");
}

static void errors(std::string const& dir, std::string const& path) {
  io::BinaryFile file(path, io::read_only);
  Cell c4(4), c3(3);
  REQUIRE(throws_with([&] { file.read_to(0, fzn_, c4); }, "Could not find chunk"));
  REQUIRE(throws_with([&] { file.read_to(0, r_, c3); }, "expected 3 rows found 4"));
  REQUIRE(throws_with([&] { (void)file.read<float>(0, "particles/N"); }, "expecting to read type 9 but found type 3"));
  system::SoA<Delta> d(8);
  REQUIRE(throws_with([&] { file.read_to(0, del_, d); }, "Could not find chunk"));
  REQUIRE(throws_with([&] { (void)file.read<std::int64_t>(0, "particles/types"); }, "expected 1 rows found 2"));
  REQUIRE(throws_with([&] { file.commit([&] { file.write("x", 1.0); }); }, "File must be writeable"));
  {
    std::ofstream junk(dir + "/junk.gsd", std::ios::binary);
    junk << std::string(300, 'x');
  }
  REQUIRE(throws_with([&] { io::BinaryFile bad(dir + "/junk.gsd"); }, "Not a GSD file"));
  REQUIRE(throws_with([&] { io::BinaryFile none(dir + "/absent/none.gsd"); }, "cannot open"));
  // Columns: a 2-column chunk (written through the file layer) read as a 3-vector property.
  {
    std::array<double, 8> v{};
    {
      io::gsd::File g(dir + "/cols.gsd", io::gsd::Mode::create, "openFLY", "hoomd", io::gsd::make_version(1, 4));
      g.write_chunk("log/particles/delta", io::gsd::Kind::f64, 4, 2, v.data());
      g.end_frame();
    }
    io::BinaryFile r(dir + "/cols.gsd");
    system::SoA<Delta> d4(4);
    REQUIRE(throws_with([&] { r.read_to(0, del_, d4); }, "expected 3 columns found 2"));
  }
  // The same chunk twice in one frame is refused; an empty chunk (N = 0) is a valid chunk.
  {
    io::BinaryFile w(dir + "/twice.gsd", io::create);
    REQUIRE(throws_with([&] { w.write("a", 1.0); w.write("a", 2.0); }, "written twice"));
    w.write(del_, system::SoA<Delta>(0));
    w.commit();
    REQUIRE(w.n_frames() == 1 && w.read<double>(0, "a") == 1.0);
    system::SoA<Delta> d0(0), d1(1);
    w.read_to(0, del_, d0);
    REQUIRE(throws_with([&] { w.read_to(0, del_, d1); }, "expected 1 rows found 0"));
  }
  // A frame whose commit never came is not in the file.
  {
    io::BinaryFile w(dir + "/uncommitted.gsd", io::create);
    w.commit([&] { w.write("v", 1.0); });
    w.write("v", 2.0);
  }
  {
    io::BinaryFile r(dir + "/uncommitted.gsd");
    REQUIRE(r.n_frames() == 1);
    REQUIRE(r.read<double>(0, "v") == 1.0);
  }
  std::printf("error paths ok\n");
}

// Many frames and names: the index (128 entries) and name list (1024 bytes) are outgrown and moved.
static void growth(std::string const& dir) {
  std::string const path = dir + "/growth.gsd";
  auto name = [](int k) { return "log/value/a_fairly_long_chunk_name_" + std::to_string(k); };
  {
    io::BinaryFile w(path, io::create);
    for (int f = 0; f < 150; ++f) {
      w.commit([&] {
        w.write("frame", static_cast<std::int32_t>(f));
        w.write(name(f % 60).c_str(), 0.5 * f);
      });
    }
  }
  auto check = [&](io::BinaryFile const& r, int frames) {
    REQUIRE(r.n_frames() == static_cast<std::uint64_t>(frames));
    for (int f = 0; f < frames; ++f) {
      REQUIRE(r.read<std::int32_t>(static_cast<std::uint64_t>(f), "frame") == f);
      REQUIRE(r.read<double>(static_cast<std::uint64_t>(f), name(f % 60).c_str()) == 0.5 * f);
    }
  };
  {
    io::BinaryFile r(path);
    check(r, 150);
    REQUIRE(r.file().header().index_allocated_entries >= 300);
    REQUIRE(r.file().header().index_location > 256);
    REQUIRE(r.file().header().namelist_allocated_entries > 1024);
  }
  {  // Reopen for writing and append.
    io::BinaryFile w(path, io::read_write);
    REQUIRE(w.n_frames() == 150);
    for (int f = 150; f < 400; ++f) {
      w.commit([&] {
        w.write("frame", static_cast<std::int32_t>(f));
        w.write(name(f % 60).c_str(), 0.5 * f);
      });
    }
  }
  {
    io::BinaryFile r(path);
    check(r, 400);
  }
  std::printf("growth ok (400 frames, 61 names, index and name list moved)\n");
}

static void clear_and_values(std::string const& dir) {
  std::string const path = dir + "/clear.gsd";
  Mat B = Mat::Zero();
  B(0, 0) = 17.13;
  B(1, 1) = 15.2;
  B(2, 2) = 14.9;
  B(0, 1) = 2.5;   // triclinic: basis columns (17.13,0,0), (2.5,15.2,0), (-1.25,3.3,14.9)
  B(0, 2) = -1.25;
  B(1, 2) = 3.3;
  system::Box box(B, Arr<bool>(true, false, true));
  system::TypeMap<Mass> masses(3);
  masses.set(0, "Fe", 55.845);
  masses.set(1, "H", 1.00784);
  masses.set(2, "Cr", 51.9961);
  {
    io::BinaryFile w(path, io::create);
    w.commit([&] { w.write("junk", 3); });
    w.clear();
    REQUIRE(w.n_frames() == 0);
    REQUIRE(throws_with([&] { (void)w.read<int>(0, "junk"); }, "Could not find chunk"));
    Cell c(3);
    for (int i = 0; i < 3; ++i) c(fzn_, i) = i != 1;
    w.commit([&] {
      w.write(box);
      w.write(masses);
      w.write(fzn_, c);
    });
  }
  auto b = slurp(path);
  REQUIRE(std::string(reinterpret_cast<char const*>(b.data() + 48)) == "openFLY");
  io::BinaryFile r(path);
  REQUIRE(r.n_frames() == 1);
  auto m = r.read_map<Mass>(0);
  REQUIRE(m.num_types() == 3);
  REQUIRE(m.get(2) == "Cr");
  REQUIRE(m.get(0, Mass{}) == 55.845 && m.get(1, Mass{}) == 1.00784 && m.get(2, Mass{}) == 51.9961);
  system::Box rb = r.read_box(0);
  for (int rr = 0; rr < 3; ++rr)
    for (int cc = 0; cc < 3; ++cc) {
      double const want = rr > cc ? 0.0 : static_cast<double>(static_cast<float>(B(rr, cc)));
      REQUIRE(rb.basis()(rr, cc) == want);
    }
  REQUIRE(rb.periodic(0) && !rb.periodic(1) && rb.periodic(2));
  Cell c(3);
  r.read_to(0, fzn_, c);
  REQUIRE(c(fzn_, 0) == 1 && c(fzn_, 1) == 0 && c(fzn_, 2) == 1);
  REQUIRE(throws_with([&] { masses.set(0, std::string(32, 'x'), 1.0); }, "must be < 32"));
  std::printf("clear, per-species values, triclinic box ok\n");
}

// The file a reader written from the published format checks: one frame of known chunks.
static void layout_file(std::string const& path) {
  io::BinaryFile w(path, io::create);
  system::SoA<Position, TypeID> s(3);
  for (int i = 0; i < 3; ++i) {
    s(r_, i) = Vec(0.25 * i, -1.5 * i, 3.0 + i);
    s(id_, i) = static_cast<std::uint32_t>(2 - i);
  }
  w.commit([&] {
    w.write("particles/N", std::uint32_t{3});
    w.write(r_, s);
    w.write(id_, s);
  });
  w.commit([&] { w.write("log/energy", -123.456); });
}

// `gsd_check --read <file>`: prints what io::BinaryFile reads from a file another writer made: the frame
// count, then per frame particles/N, the positions (frame 0's when a frame has none) and log/value.
static int read_foreign(char const* path) {
  try {
    io::BinaryFile f(path);
    std::printf("frames %llu\n", static_cast<unsigned long long>(f.n_frames()));
    for (std::uint64_t i = 0; i < f.n_frames(); ++i) {
      auto const n = f.read<std::uint32_t>(i, "particles/N");
      system::SoA<Position> s(n);
      f.read_to(i, r_, s);
      std::printf("frame %llu N %u value %d pos", static_cast<unsigned long long>(i), n, f.read<std::int32_t>(i, "log/value"));
      for (std::size_t k = 0; k < s[r_].size(); ++k) std::printf(" %.17g", s[r_][k]);
      std::printf("\n");
    }
  } catch (std::exception const& e) {
    std::printf("error %s\n", e.what());
    return 1;
  }
  return 0;
}

int main(int argc, char** argv) {
  if (argc == 3 && std::string(argv[1]) == "--read") return read_foreign(argv[2]);
  if (argc < 2) {
    std::fprintf(stderr, "usage: gsd_check <directory> | gsd_check --read <file>\n");
    return 2;
  }
  std::string const dir = argv[1];
  std::string const path = dir + "/BinaryFile_test.gsd";
  try {
    reference_round_trip(path);
    header_layout(path);
    errors(dir, path);
    growth(dir);
    clear_and_values(dir);
    layout_file(dir + "/layout.gsd");
  } catch (std::exception const& e) {
    std::fprintf(stderr, "exception: %s\n", e.what());
    return 1;
  }
  if (failures) {
    std::fprintf(stderr, "%d failures\n", failures);
    return 1;
  }
  std::printf("gsd ok\n");
  return 0;
}
