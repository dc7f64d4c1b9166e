// Reader of WSClean's reordered measurement-set part files
// (msproviders/reordering.h:22-60, msproviders/reorderedmsprovider.cpp:58-97,
// msproviders/msreaders/reorderedmsreader.cpp:6-150): the per-spectral-window
// meta file and one part's data, weight and model files.  Host code only.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>

namespace wg {

// Packed sizes of the reordered files' records (reordering.h:22-44).
constexpr size_t kMetaHeaderBytes = 20;  // start_time f64, rows u64, filename length u32
constexpr size_t kMetaRecordBytes = 38;  // u, v, w, time f64; antenna1, antenna2, field u16
constexpr size_t kPartHeaderBytes = 21;  // channel count u64, start u64, data desc u32, model u8

// One reordered part: rows of the meta file (shared by the parts of a
// spectral window) with the part's channel range of visibilities, weights
// and, when the part header says so, model visibilities.  n_pol is the
// number of visibilities per row and channel in the files (1 for a single
// requested polarization).  Reads use pread at computed offsets, so the
// reader keeps no file position and reads of disjoint rows may run in
// parallel.
class PartReader {
 public:
  PartReader(const std::string& meta_path, const std::string& data_path,
             const std::string& weight_path, const std::string& model_path, size_t n_pol);
  ~PartReader();
  PartReader(const PartReader&) = delete;
  PartReader& operator=(const PartReader&) = delete;

  size_t Rows() const { return rows_; }
  size_t Channels() const { return n_chan_; }
  size_t ChannelStart() const { return chan_start_; }
  uint32_t DataDescId() const { return data_desc_id_; }
  bool HasModel() const { return has_model_; }
  double StartTime() const { return start_time_; }
  size_t Polarizations() const { return n_pol_; }
  const std::string& MsPath() const { return ms_path_; }
  bool HasWeights() const { return fd_weight_ >= 0; }

  // Rows [row0, row0 + n): meta records unpacked into uvw [n][3] (metres)
  // and, where the pointers are non-null, antenna pairs [n][2], times [n]
  // and field ids [n].
  void ReadMeta(size_t row0, size_t n, double* uvw, uint16_t* antennas, double* times,
                uint16_t* fields) const;
  // The raw packed records of rows [row0, row0 + n) (38 bytes each).
  void ReadMetaRaw(size_t row0, size_t n, void* records) const;
  // vis [n][n_chan][n_pol] complex<float>
  void ReadData(size_t row0, size_t n, float* vis) const;
  // weights [n][n_chan][n_pol] float (1 when the part has no weight file)
  void ReadWeights(size_t row0, size_t n, float* weights) const;
  // model [n][n_chan][n_pol] complex<float> (the part must have a model)
  void ReadModel(size_t row0, size_t n, float* model) const;

  // Unpacks n raw meta records into uvw [n][3].
  static void UnpackUvw(const void* records, size_t n, double* uvw);

 private:
  void ReadAt(int fd, void* dst, size_t bytes, size_t offset, const char* what) const;
  int fd_meta_ = -1, fd_data_ = -1, fd_weight_ = -1, fd_model_ = -1;
  size_t rows_ = 0, n_chan_ = 0, chan_start_ = 0, n_pol_ = 1, meta_offset_ = 0;
  uint32_t data_desc_id_ = 0;
  bool has_model_ = false;
  double start_time_ = 0;
  std::string ms_path_;
};

}  // namespace wg
