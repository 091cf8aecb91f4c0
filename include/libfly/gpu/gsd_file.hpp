// GSD trajectory files for the C++ drop-in: fly::io::BinaryFile with the reference's member signatures
// (gsd.hpp:118-327, gsd.cpp:63-253), over a file layer written here from the published GSD 2.0 file
// layout (the gsd C library that the reference links is not used).  Host code only, POSIX I/O.
//
// File layout (GSD 2.0):
//   header (256 bytes at offset 0): magic 0x65DF65DF65DF65DF, index location and allocated entries,
//     name-list location and allocated bytes, schema version, file version (2.0), application[64],
//     schema[64], reserved[80];
//   index: an array of 32-byte entries (frame, N, location, M, id, type, flags), ordered by frame and,
//     inside a frame, by name id; unused entries have location 0.  When full, a twice larger index is
//     written at the end of the file and the header points to it;
//   name list: the chunk names as consecutive NUL-terminated strings (name id = position in the list),
//     zero padded to its allocated size, which doubles (moved to the end of the file) when full;
//   chunk data: N x M values of one type each, appended to the end of the file as they are written.
// A frame is written as chunks followed by commit(); its index entries (and new names) reach the file
// at commit, so a frame not committed when the file is closed is not part of the file.
// Semantics of the frame data (gsd.hpp:36-52): a chunk absent from frame i reads as the chunk of the
// same name at frame 0.
//
// Included by facade.hpp after the system types (Box, TypeMap, SoA, properties) it reads and writes, and
// before the minimiser and dimer whose Options::fout hooks write frames; include libfly/gpu/gsd.hpp or
// facade.hpp, not this file.
#pragma once

#ifndef FLY_GPU_FACADE_SYSTEM_TYPES
#error "include libfly/gpu/gsd.hpp or libfly/gpu/facade.hpp"
#endif

#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <array>
#include <cerrno>
#include <cstdint>
#include <cstring>
#include <functional>
#include <string>
#include <string_view>
#include <type_traits>
#include <unordered_map>
#include <utility>
#include <vector>

namespace fly::io {

  namespace gsd {

    constexpr std::uint64_t magic = 0x65DF65DF65DF65DFull;

    constexpr std::uint32_t make_version(std::uint32_t major, std::uint32_t minor) { return (major << 16) | minor; }

    // Element types of a chunk and their sizes.
    enum class Kind : std::uint8_t { u8 = 1, u16, u32, u64, i8, i16, i32, i64, f32, f64 };

    inline std::size_t size_of(std::uint8_t kind) {
      switch (kind) {
        case 1: case 5: return 1;
        case 2: case 6: return 2;
        case 3: case 7: case 9: return 4;
        case 4: case 8: case 10: return 8;
        default: return 0;
      }
    }

    template <typename T>
    constexpr Kind kind_of() {
      static_assert(std::is_arithmetic_v<T>, "GSD chunks hold arithmetic values");
      if constexpr (std::is_floating_point_v<T>) {
        static_assert(sizeof(T) == 4 || sizeof(T) == 8, "no GSD floating point type of this size");
        return sizeof(T) == 4 ? Kind::f32 : Kind::f64;
      } else if constexpr (std::is_signed_v<T>) {
        static_assert(sizeof(T) == 1 || sizeof(T) == 2 || sizeof(T) == 4 || sizeof(T) == 8, "no GSD signed type of this size");
        return sizeof(T) == 1 ? Kind::i8 : sizeof(T) == 2 ? Kind::i16 : sizeof(T) == 4 ? Kind::i32 : Kind::i64;
      } else {
        static_assert(sizeof(T) == 1 || sizeof(T) == 2 || sizeof(T) == 4 || sizeof(T) == 8, "no GSD unsigned type of this size");
        return sizeof(T) == 1 ? Kind::u8 : sizeof(T) == 2 ? Kind::u16 : sizeof(T) == 4 ? Kind::u32 : Kind::u64;
      }
    }

    struct Header {
      std::uint64_t magic;
      std::uint64_t index_location;
      std::uint64_t index_allocated_entries;
      std::uint64_t namelist_location;
      std::uint64_t namelist_allocated_entries;   // bytes of the name list
      std::uint32_t schema_version;
      std::uint32_t gsd_version;
      char application[64];
      char schema[64];
      char reserved[80];
    };
    static_assert(sizeof(Header) == 256, "GSD header is 256 bytes");

    struct IndexEntry {
      std::uint64_t frame;
      std::uint64_t N;
      std::int64_t location;
      std::uint32_t M;
      std::uint16_t id;
      std::uint8_t type;
      std::uint8_t flags;
    };
    static_assert(sizeof(IndexEntry) == 32, "GSD index entry is 32 bytes");

    constexpr std::uint64_t initial_index_entries = 128;
    constexpr std::uint64_t initial_namelist_bytes = 1024;

    enum class Mode { read_only, read_write, create };

    // One open GSD file.
    class File {
    public:
      File(std::string path, Mode mode, char const* application = "", char const* schema = "",
           std::uint32_t schema_version = 0)
          : m_path(std::move(path)), m_writable(mode != Mode::read_only) {
        int const flags = mode == Mode::read_only ? O_RDONLY : mode == Mode::read_write ? O_RDWR : (O_RDWR | O_CREAT | O_TRUNC);
        m_fd = ::open(m_path.c_str(), flags, 0644);
        if (m_fd < 0) fail("cannot open file (" + std::string(std::strerror(errno)) + ")");
        try {
          if (mode == Mode::create) {
            Header h{};
            h.magic = magic;
            h.gsd_version = make_version(2, 0);
            h.schema_version = schema_version;
            std::strncpy(h.application, application, sizeof(h.application) - 1);
            std::strncpy(h.schema, schema, sizeof(h.schema) - 1);
            initialise(h);
          } else {
            load();
          }
        } catch (...) {
          ::close(m_fd);
          throw;
        }
      }
      File(File const&) = delete;
      File& operator=(File const&) = delete;
      ~File() {
        if (m_fd >= 0) ::close(m_fd);
      }

      Header const& header() const { return m_h; }
      std::uint64_t n_frames() const { return m_frame; }

      // Empties the file (no frames, no names), keeping application, schema and schema version.
      void truncate() {
        require_writable();
        if (::ftruncate(m_fd, 0) != 0) fail("cannot truncate (" + std::string(std::strerror(errno)) + ")");
        Header h = m_h;
        initialise(h);
      }

      void write_chunk(char const* name, Kind kind, std::uint64_t N, std::uint32_t M, void const* data) {
        require_writable();
        std::string_view const nm{name};
        if (nm.empty()) fail("empty chunk name");
        if (M == 0) fail("chunk '" + std::string(nm) + "' has zero columns");
        std::uint16_t const id = name_id(nm, true);
        for (auto const& e : m_pending)
          if (e.id == id) fail("chunk '" + std::string(nm) + "' written twice in frame " + std::to_string(m_frame));
        std::uint64_t const bytes = N * M * size_of(static_cast<std::uint8_t>(kind));
        IndexEntry e{m_frame, N, static_cast<std::int64_t>(m_end), M, id, static_cast<std::uint8_t>(kind), 0};
        if (bytes) write_at(m_end, data, bytes);
        m_end += bytes;
        m_pending.push_back(e);
      }

      // Commits the chunks written since the last commit as the next frame.
      void end_frame() {
        require_writable();
        flush_names();
        std::sort(m_pending.begin(), m_pending.end(), [](IndexEntry const& a, IndexEntry const& b) { return a.id < b.id; });
        std::uint64_t const need = m_index.size() + m_pending.size();
        if (need > m_h.index_allocated_entries) {
          std::uint64_t cap = m_h.index_allocated_entries;
          while (cap < need) cap *= 2;
          // New index at the end of the file: the committed entries, then zeros.
          std::vector<IndexEntry> block(cap, IndexEntry{});
          std::copy(m_index.begin(), m_index.end(), block.begin());
          std::uint64_t const at = m_end;
          write_at(at, block.data(), cap * sizeof(IndexEntry));
          m_end += cap * sizeof(IndexEntry);
          m_h.index_location = at;
          m_h.index_allocated_entries = cap;
          write_header();
        }
        if (!m_pending.empty()) {
          write_at(m_h.index_location + m_index.size() * sizeof(IndexEntry), m_pending.data(),
                   m_pending.size() * sizeof(IndexEntry));
        }
        m_index.insert(m_index.end(), m_pending.begin(), m_pending.end());
        m_pending.clear();
        ++m_frame;
      }

      // The entry of chunk `name` in `frame`, or null when that frame holds no such chunk.
      IndexEntry const* find(std::uint64_t frame, std::string_view name) const {
        if (frame >= m_frame) return nullptr;
        auto const it = m_ids.find(std::string(name));
        if (it == m_ids.end() || it->second >= m_names_on_disk) return nullptr;
        auto const lo = std::lower_bound(m_index.begin(), m_index.end(), std::make_pair(frame, it->second),
                                         [](IndexEntry const& e, std::pair<std::uint64_t, std::uint16_t> const& k) {
                                           return e.frame < k.first || (e.frame == k.first && e.id < k.second);
                                         });
        if (lo == m_index.end() || lo->frame != frame || lo->id != it->second) return nullptr;
        return &*lo;
      }

      void read(IndexEntry const& e, void* out) const {
        std::uint64_t const bytes = e.N * e.M * size_of(e.type);
        if (bytes) read_at(static_cast<std::uint64_t>(e.location), out, bytes);
      }

      std::string const& path() const { return m_path; }

    private:
      [[noreturn]] void fail(std::string const& what) const { throw RuntimeError("GSD: " + what + " - " + m_path); }

      void require_writable() const {
        if (!m_writable) fail("File must be writeable");
      }

      void write_at(std::uint64_t off, void const* p, std::uint64_t n) {
        auto const* c = static_cast<char const*>(p);
        while (n) {
          ssize_t const w = ::pwrite(m_fd, c, n, static_cast<off_t>(off));
          if (w < 0 && errno == EINTR) continue;
          if (w <= 0) fail("write failed (" + std::string(std::strerror(errno)) + ")");
          c += w;
          off += static_cast<std::uint64_t>(w);
          n -= static_cast<std::uint64_t>(w);
        }
      }

      void read_at(std::uint64_t off, void* p, std::uint64_t n) const {
        auto* c = static_cast<char*>(p);
        while (n) {
          ssize_t const r = ::pread(m_fd, c, n, static_cast<off_t>(off));
          if (r < 0 && errno == EINTR) continue;
          if (r <= 0) fail("read past the end of the file or failed");
          c += r;
          off += static_cast<std::uint64_t>(r);
          n -= static_cast<std::uint64_t>(r);
        }
      }

      void write_header() { write_at(0, &m_h, sizeof(Header)); }

      // A fresh file: header, an empty index, an empty name list.
      void initialise(Header h) {
        h.index_location = sizeof(Header);
        h.index_allocated_entries = initial_index_entries;
        h.namelist_location = h.index_location + initial_index_entries * sizeof(IndexEntry);
        h.namelist_allocated_entries = initial_namelist_bytes;
        std::memset(h.reserved, 0, sizeof(h.reserved));
        m_h = h;
        write_header();
        std::vector<char> zeros(initial_index_entries * sizeof(IndexEntry) + initial_namelist_bytes, 0);
        write_at(sizeof(Header), zeros.data(), zeros.size());
        m_end = h.namelist_location + initial_namelist_bytes;
        m_index.clear();
        m_pending.clear();
        m_names.clear();
        m_ids.clear();
        m_names_on_disk = 0;
        m_names_bytes = 0;
        m_frame = 0;
      }

      void load() {
        struct stat st {};
        if (::fstat(m_fd, &st) != 0) fail("cannot stat");
        auto const size = static_cast<std::uint64_t>(st.st_size);
        if (size < sizeof(Header)) fail("Not a GSD file");
        read_at(0, &m_h, sizeof(Header));
        if (m_h.magic != magic) fail("Not a GSD file");
        if ((m_h.gsd_version >> 16) != 2) fail("Invalid GSD file version " + std::to_string(m_h.gsd_version >> 16) + "." +
                                                std::to_string(m_h.gsd_version & 0xffff));
        if (m_h.index_location + m_h.index_allocated_entries * sizeof(IndexEntry) > size ||
            m_h.namelist_location + m_h.namelist_allocated_entries > size) {
          fail("File corrupt (index or name list past the end of the file)");
        }
        // Names: NUL-terminated strings up to the first empty one.
        std::vector<char> nl(m_h.namelist_allocated_entries);
        if (!nl.empty()) read_at(m_h.namelist_location, nl.data(), nl.size());
        std::size_t p = 0;
        while (p < nl.size() && nl[p] != 0) {
          std::size_t q = p;
          while (q < nl.size() && nl[q] != 0) ++q;
          if (q == nl.size()) fail("File corrupt (unterminated name)");
          m_ids.emplace(std::string(nl.data() + p, q - p), static_cast<std::uint16_t>(m_names.size()));
          m_names.emplace_back(nl.data() + p, q - p);
          p = q + 1;
        }
        m_names_on_disk = m_names.size();
        m_names_bytes = p;
        // Index: the used entries (location != 0) form a prefix; looked up by (frame, id).
        std::vector<IndexEntry> idx(m_h.index_allocated_entries);
        if (!idx.empty()) read_at(m_h.index_location, idx.data(), idx.size() * sizeof(IndexEntry));
        std::size_t used = 0;
        while (used < idx.size() && idx[used].location != 0) ++used;
        idx.resize(used);
        std::stable_sort(idx.begin(), idx.end(), [](IndexEntry const& a, IndexEntry const& b) {
          return a.frame < b.frame || (a.frame == b.frame && a.id < b.id);
        });
        for (std::size_t i = 0; i < used; ++i) {
          IndexEntry const& e = idx[i];
          if (e.id >= m_names.size() || size_of(e.type) == 0 || e.location < 0 ||
              static_cast<std::uint64_t>(e.location) + e.N * e.M * size_of(e.type) > size ||
              (i && e.frame == idx[i - 1].frame && e.id == idx[i - 1].id)) {
            fail("File corrupt (index entry " + std::to_string(i) + ")");
          }
        }
        m_index = std::move(idx);
        m_frame = m_index.empty() ? 0 : m_index.back().frame + 1;
        m_end = size;
      }

      std::uint16_t name_id(std::string_view name, bool add) {
        auto const it = m_ids.find(std::string(name));
        if (it != m_ids.end()) return it->second;
        if (!add) fail("unknown chunk name");
        if (m_names.size() >= 0xffff) fail("too many chunk names");
        auto const id = static_cast<std::uint16_t>(m_names.size());
        m_names.emplace_back(name);
        m_ids.emplace(std::string(name), id);
        return id;
      }

      // Writes the names added since the last commit to the name list, moving it to a larger block at
      // the end of the file when they do not fit (one zero byte always ends the list).
      void flush_names() {
        if (m_names_on_disk == m_names.size()) return;
        std::string add;
        for (std::size_t i = m_names_on_disk; i < m_names.size(); ++i) {
          add += m_names[i];
          add.push_back('\0');
        }
        std::uint64_t const need = m_names_bytes + add.size() + 1;
        if (need > m_h.namelist_allocated_entries) {
          std::uint64_t cap = m_h.namelist_allocated_entries;
          while (cap < need) cap *= 2;
          std::vector<char> block(cap, 0);
          std::size_t p = 0;
          for (std::size_t i = 0; i < m_names_on_disk; ++i) {
            std::memcpy(block.data() + p, m_names[i].data(), m_names[i].size());
            p += m_names[i].size() + 1;
          }
          std::memcpy(block.data() + p, add.data(), add.size());
          std::uint64_t const at = m_end;
          write_at(at, block.data(), cap);
          m_end += cap;
          m_h.namelist_location = at;
          m_h.namelist_allocated_entries = cap;
          write_header();
        } else {
          write_at(m_h.namelist_location + m_names_bytes, add.data(), add.size());
        }
        m_names_bytes += add.size();
        m_names_on_disk = m_names.size();
      }

      std::string m_path;
      bool m_writable;
      int m_fd = -1;
      Header m_h{};
      std::vector<IndexEntry> m_index;     // committed entries, as in the file
      std::vector<IndexEntry> m_pending;   // entries of the frame being written
      std::vector<std::string> m_names;
      std::unordered_map<std::string, std::uint16_t> m_ids;
      std::size_t m_names_on_disk = 0;
      std::uint64_t m_names_bytes = 0;     // bytes of the names in the file's name list
      std::uint64_t m_end = 0;             // end of the file
      std::uint64_t m_frame = 0;           // frames committed
    };

  }  // namespace gsd

  // File permissions (gsd.hpp:103-111).
  enum Flags {
    read_only,   // read only
    read_write,  // read and write
    create       // read and write, overwriting an existing file
  };

  namespace detail {
    // "log/typemap/<tag without a leading log/>" (gsd.hpp:63-99): per-species properties of a TypeMap get
    // their own labels, so readers that treat "log/particles/*" as per-particle data are not misled.
    inline std::string typemap_tag(char const* tag) {
      std::string_view t{tag};
      if (t.size() >= 4 && t.substr(0, 4) == "log/") t.remove_prefix(4);
      return "log/typemap/" + std::string(t);
    }
  }  // namespace detail

  // A GSD file (gsd.hpp:113-327): frames of chunks; the box in HOOMD's schema, TypeMaps and SoA
  // properties under their chunk labels.
  class BinaryFile {
    struct noop {
      constexpr void operator()() const noexcept {}
    };

  public:
    // Creating a file writes application "openFLY", schema "hoomd", schema version 1.4 (gsd.cpp:63-80).
    explicit BinaryFile(std::string_view fname, Flags flag = read_only)
        : m_file(std::string(fname),
                 flag == read_only ? gsd::Mode::read_only : flag == read_write ? gsd::Mode::read_write : gsd::Mode::create,
                 "openFLY", "hoomd", gsd::make_version(1, 4)) {}

    auto n_frames() const noexcept -> std::uint64_t { return m_file.n_frames(); }

    // No frames and no chunks afterwards; application and schema are kept.
    auto clear() -> void { m_file.truncate(); }

    // Calls f (which writes the frame's chunks), then commits the frame.
    template <typename F = noop>
    auto commit(F&& f = {}) -> void {
      std::invoke(std::forward<F>(f));
      m_file.end_frame();
    }

    template <typename T>
    auto write(char const* name, T const& value) -> std::enable_if_t<std::is_arithmetic_v<T>> {
      write_chunk(name, 1, 1, &value);
    }

    // HOOMD's box chunk, six single precision floats (basis entries (0,0), (1,1), (2,2), (0,1), (0,2),
    // (1,2), as the reference stores them: lossy), and the periodicity as three bytes (gsd.cpp:96-127).
    auto write(system::Box const& box) -> void {
      double const* b = box.basis_data();   // row-major, columns are the basis vectors
      std::array<float, 6> const hoomd{static_cast<float>(b[0]), static_cast<float>(b[4]), static_cast<float>(b[8]),
                                       static_cast<float>(b[1]), static_cast<float>(b[2]), static_cast<float>(b[5])};
      write_chunk("configuration/box", 6, 1, hoomd.data());
      std::array<std::uint8_t, 3> const per{box.periodic(0), box.periodic(1), box.periodic(2)};
      write_chunk("log/periodicity", 3, 1, per.data());
    }

    // The species count, names (num_types x 32 characters) and per-species values (gsd.hpp:196-206).
    template <typename... U>
    auto write(system::TypeMap<U...> const& map) -> void {
      write("log/typemap/N", static_cast<std::int64_t>(map.num_types()));
      std::vector<char> names(static_cast<std::size_t>(map.num_types()) * Type::size(), 0);
      for (int i = 0; i < map.num_types(); ++i) {
        std::string const& s = map.get(i);
        std::memcpy(names.data() + static_cast<std::size_t>(i) * Type::size(), s.data(), s.size());
      }
      write_chunk(Type::tag, map.num_types(), Type::size(), names.data());
      (write_chunk(detail::typemap_tag(U::tag).c_str(), map.num_types(), U::size(), map.values()[U{}].data()), ...);
    }

    // One property of an SoA: size() x T::size() values (gsd.hpp:208-219).
    template <typename T, typename... U>
    auto write(T, system::SoA<U...> const& in) -> std::enable_if_t<std::is_arithmetic_v<typename T::scalar_t>> {
      write_chunk(T::tag, in.size(), T::size(), in[T{}].data());
    }

    template <typename T>
    auto read(std::uint64_t i, char const* name) const -> std::enable_if_t<std::is_arithmetic_v<T>, T> {
      T v{};
      read_chunk(i, name, 1, 1, &v);
      return v;
    }

    // The box of frame i: the six floats widened to double, the lower triangle zero (gsd.cpp:129-173).
    auto read_box(std::uint64_t i) const -> system::Box {
      std::array<float, 6> h{};
      read_chunk(i, "configuration/box", 6, 1, h.data());
      std::array<std::uint8_t, 3> per{};
      read_chunk(i, "log/periodicity", 3, 1, per.data());
      Mat m = Mat::Zero();
      m(0, 0) = h[0];
      m(1, 1) = h[1];
      m(2, 2) = h[2];
      m(0, 1) = h[3];
      m(0, 2) = h[4];
      m(1, 2) = h[5];
      return {m, Arr<bool>(per[0] != 0, per[1] != 0, per[2] != 0)};
    }

    // Reads property T of frame i into out, whose size must match the chunk's rows (gsd.hpp:239-253).
    template <typename T, typename... U>
    auto read_to(std::uint64_t i, T, system::SoA<U...>& out) const -> std::enable_if_t<std::is_arithmetic_v<typename T::scalar_t>> {
      read_chunk(i, T::tag, out.size(), T::size(), out[T{}].data());
    }

    template <typename... U>
    auto read_map(std::uint64_t i) const -> system::TypeMap<U...> {
      auto const n = read<std::int64_t>(i, "log/typemap/N");
      if (n < 0 || n > (1 << 20)) throw RuntimeError("GSD: implausible species count " + std::to_string(n));
      system::TypeMap<U...> map(static_cast<int>(n));
      std::vector<char> names(static_cast<std::size_t>(n) * Type::size());
      read_chunk(i, Type::tag, n, Type::size(), names.data());
      for (int t = 0; t < n; ++t) {
        char const* p = names.data() + static_cast<std::size_t>(t) * Type::size();
        std::size_t len = 0;
        while (len < static_cast<std::size_t>(Type::size()) && p[len]) ++len;
        if (len == static_cast<std::size_t>(Type::size())) throw RuntimeError("GSD: species name not terminated");
        map.set_name(t, std::string(p, len));
      }
      (read_chunk(i, detail::typemap_tag(U::tag).c_str(), n, U::size(), map.values()[U{}].data()), ...);
      return map;
    }

    // The file layer, for inspection.
    gsd::File const& file() const { return m_file; }

  private:
    template <typename T>
    void write_chunk(char const* name, std::int64_t N, int M, T const* data) {
      if (N < 0 || M <= 0) throw RuntimeError(std::string("GSD: bad shape for chunk '") + name + "'");
      m_file.write_chunk(name, gsd::kind_of<T>(), static_cast<std::uint64_t>(N), static_cast<std::uint32_t>(M), data);
    }

    // Frame i's chunk, or frame 0's when frame i has none (gsd.cpp:221-251); shape and type must match.
    template <typename T>
    void read_chunk(std::uint64_t frame, char const* name, std::int64_t N, int M, T* data) const {
      gsd::IndexEntry const* e = m_file.find(frame, name);
      if (!e) frame = 0;
      if (!e) e = m_file.find(0, name);
      if (!e) throw RuntimeError(std::string("GSD: Could not find chunk with name '") + name + "' at frame 0");
      std::string const where = std::string("GSD: Chunk '") + name + "' frame " + std::to_string(frame);
      if (static_cast<std::uint64_t>(N) != e->N) {
        throw RuntimeError(where + ", expected " + std::to_string(N) + " rows found " + std::to_string(e->N));
      }
      if (static_cast<std::uint32_t>(M) != e->M) {
        throw RuntimeError(where + ", expected " + std::to_string(M) + " columns found " + std::to_string(e->M));
      }
      auto const k = static_cast<std::uint8_t>(gsd::kind_of<T>());
      if (e->type != k) {
        throw RuntimeError(where + ", expecting to read type " + std::to_string(k) + " but found type " + std::to_string(e->type));
      }
      m_file.read(*e, data);
    }

    gsd::File m_file;
  };

}  // namespace fly::io
