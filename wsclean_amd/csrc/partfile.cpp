// Reader of WSClean's reordered part files (host code).  Layouts, all
// little-endian and packed (msproviders/reordering.h:22-60):
//   meta file  <prefix>-spw<d>-parted-meta.tmp: MetaHeader {start_time f64,
//              selected_row_count u64, filename_length u32} (20 bytes), the
//              measurement-set path (filename_length bytes), then one
//              MetaRecord {u, v, w, time f64; antenna1, antenna2, field_id
//              u16} (38 bytes) per selected row;
//   data file  <part prefix>.tmp: PartHeader {channel_count u64,
//              channel_start u64, data_desc_id u32, has_model u8} (21 bytes),
//              then per row channel_count x n_pol complex<float>
//              (msproviders/msreaders/reorderedmsreader.cpp:21-34, :96-126);
//   weights    <part prefix>-w.tmp: per row channel_count x n_pol float, no
//              header (reorderedmsreader.cpp:35-39, :140-155);
//   model      <part prefix>-m.tmp: as the data without the header
//              (msproviders/reorderedmsprovider.cpp:88-93).
#include "partfile.hpp"

#include <fcntl.h>
#include <unistd.h>

#include <cerrno>
#include <cstring>
#include <vector>

#include "common.hpp"

namespace wg {

namespace {

int OpenRead(const std::string& path, const char* what) {
  const int fd = ::open(path.c_str(), O_RDONLY | O_CLOEXEC);
  WG_CHECK(fd >= 0, std::string("cannot open the ") + what + " file '" + path + "': " +
                        std::strerror(errno));
  return fd;
}

size_t FileSize(int fd) {
  const off_t end = ::lseek(fd, 0, SEEK_END);
  WG_CHECK(end >= 0, "cannot size a reordered file");
  return size_t(end);
}

template <typename T>
T Field(const unsigned char* p) {
  T v;
  std::memcpy(&v, p, sizeof(T));
  return v;
}

}  // namespace

void PartReader::ReadAt(int fd, void* dst, size_t bytes, size_t offset, const char* what) const {
  unsigned char* d = static_cast<unsigned char*>(dst);
  while (bytes > 0) {
    const ssize_t r = ::pread(fd, d, bytes, off_t(offset));
    if (r < 0 && errno == EINTR) continue;
    WG_CHECK(r > 0, std::string("short read of the reordered ") + what + " file");
    d += r;
    bytes -= size_t(r);
    offset += size_t(r);
  }
}

PartReader::PartReader(const std::string& meta_path, const std::string& data_path,
                       const std::string& weight_path, const std::string& model_path,
                       size_t n_pol)
    : n_pol_(n_pol) {
  WG_CHECK(n_pol == 1 || n_pol == 2 || n_pol == 4, "1, 2 or 4 polarizations per file");
  try {
    fd_meta_ = OpenRead(meta_path, "meta");
    unsigned char mh[kMetaHeaderBytes];
    ReadAt(fd_meta_, mh, sizeof(mh), 0, "meta");
    start_time_ = Field<double>(mh);
    rows_ = size_t(Field<uint64_t>(mh + 8));
    const uint32_t name_len = Field<uint32_t>(mh + 16);
    std::vector<char> name(name_len);
    if (name_len) ReadAt(fd_meta_, name.data(), name_len, kMetaHeaderBytes, "meta");
    ms_path_.assign(name.begin(), name.end());
    meta_offset_ = kMetaHeaderBytes + name_len;
    WG_CHECK(FileSize(fd_meta_) >= meta_offset_ + rows_ * kMetaRecordBytes,
             "the meta file holds fewer records than its header's row count");

    fd_data_ = OpenRead(data_path, "data");
    unsigned char ph[kPartHeaderBytes];
    ReadAt(fd_data_, ph, sizeof(ph), 0, "data");
    n_chan_ = size_t(Field<uint64_t>(ph));
    chan_start_ = size_t(Field<uint64_t>(ph + 8));
    data_desc_id_ = Field<uint32_t>(ph + 16);
    has_model_ = ph[20] != 0;
    WG_CHECK(n_chan_ > 0, "the part has no channels");
    const size_t row_vis = n_chan_ * n_pol_;
    WG_CHECK(FileSize(fd_data_) >= kPartHeaderBytes + rows_ * row_vis * 8,
             "the data file is shorter than rows x channels x polarizations");
    if (!weight_path.empty()) {
      fd_weight_ = OpenRead(weight_path, "weight");
      WG_CHECK(FileSize(fd_weight_) >= rows_ * row_vis * 4,
               "the weight file is shorter than rows x channels x polarizations");
    }
    if (has_model_ && !model_path.empty()) {
      fd_model_ = OpenRead(model_path, "model");
      WG_CHECK(FileSize(fd_model_) >= rows_ * row_vis * 8,
               "the model file is shorter than rows x channels x polarizations");
    }
  } catch (...) {
    for (int fd : {fd_meta_, fd_data_, fd_weight_, fd_model_})
      if (fd >= 0) ::close(fd);
    throw;
  }
}

PartReader::~PartReader() {
  for (int fd : {fd_meta_, fd_data_, fd_weight_, fd_model_})
    if (fd >= 0) ::close(fd);
}

void PartReader::ReadMetaRaw(size_t row0, size_t n, void* records) const {
  WG_CHECK(row0 + n <= rows_, "rows beyond the part");
  if (n) ReadAt(fd_meta_, records, n * kMetaRecordBytes, meta_offset_ + row0 * kMetaRecordBytes,
                "meta");
}

void PartReader::UnpackUvw(const void* records, size_t n, double* uvw) {
  const unsigned char* p = static_cast<const unsigned char*>(records);
  for (size_t r = 0; r < n; ++r, p += kMetaRecordBytes) {
    uvw[3 * r] = Field<double>(p);
    uvw[3 * r + 1] = Field<double>(p + 8);
    uvw[3 * r + 2] = Field<double>(p + 16);
  }
}

void PartReader::ReadMeta(size_t row0, size_t n, double* uvw, uint16_t* antennas, double* times,
                          uint16_t* fields) const {
  std::vector<unsigned char> rec(n * kMetaRecordBytes);
  ReadMetaRaw(row0, n, rec.data());
  if (uvw) UnpackUvw(rec.data(), n, uvw);
  for (size_t r = 0; r < n; ++r) {
    const unsigned char* p = rec.data() + r * kMetaRecordBytes;
    if (times) times[r] = Field<double>(p + 24);
    if (antennas) {
      antennas[2 * r] = Field<uint16_t>(p + 32);
      antennas[2 * r + 1] = Field<uint16_t>(p + 34);
    }
    if (fields) fields[r] = Field<uint16_t>(p + 36);
  }
}

void PartReader::ReadData(size_t row0, size_t n, float* vis) const {
  WG_CHECK(row0 + n <= rows_, "rows beyond the part");
  const size_t row_bytes = n_chan_ * n_pol_ * 8;
  if (n) ReadAt(fd_data_, vis, n * row_bytes, kPartHeaderBytes + row0 * row_bytes, "data");
}

void PartReader::ReadWeights(size_t row0, size_t n, float* weights) const {
  WG_CHECK(row0 + n <= rows_, "rows beyond the part");
  const size_t row_vals = n_chan_ * n_pol_;
  if (fd_weight_ < 0) {
    for (size_t i = 0; i < n * row_vals; ++i) weights[i] = 1.0f;
    return;
  }
  if (n) ReadAt(fd_weight_, weights, n * row_vals * 4, row0 * row_vals * 4, "weight");
}

void PartReader::ReadModel(size_t row0, size_t n, float* model) const {
  WG_CHECK(row0 + n <= rows_, "rows beyond the part");
  WG_CHECK(fd_model_ >= 0, "the part has no model file");
  const size_t row_bytes = n_chan_ * n_pol_ * 8;
  if (n) ReadAt(fd_model_, model, n * row_bytes, row0 * row_bytes, "model");
}

}  // namespace wg
