// poutre_amd.hpp -- C++ API of the MI355X morphology library, mirroring Poutre2's public API
// for the structuring-element morphology and geodesy path, so that host code written
// against Poutre2 compiles and runs unchanged on the GPU implementation.
//
// Mirrored declarations (reference paths relative to the Poutre2 source root):
//   poutre::PType / CompoundType           include/poutre/base/types.hpp:505-527
//   poutre::IInterface, Assert*, Create     include/poutre/base/image_interface.hpp:39-106
//   ImageFromString / ImageToString         src/base/image_interface.cpp:407-425,616-909
//   se::Common_NL_SE / Compound_NL_SE       include/poutre/structuring_element/se_types_and_tags.hpp:45-63
//   se::IStructuringElement                 include/poutre/structuring_element/se_interface.hpp:32-58
//   se::neighbor_list (runtime SE) + SESquare2D ... predefined objects
//                                           neighbor_list_se_t.hpp:36-228, predefined_nl_se.cpp:19-91
//   Erode/Dilate (3 overloads)              include/poutre/low_level_morpho/ero_dil.hpp:32-47
//   ErodeX/DilateX/ErodeY/DilateY           include/poutre/low_level_morpho/ero_dil_line.hpp:32-47
//   geo::Reconstruction                     include/poutre/geodesy/mreconstruct.hpp:30-41
//   geo::low_/high_leveling, leveling       include/poutre/geodesy/leveling.hpp:37-52
//   geo::h_minima ... dynamic_pseudo_*      include/poutre/geodesy/extrema.hpp:33-64
// Errors are std::runtime_error whose message starts with "Throw " like the reference's
// POUTRE_RUNTIME_ERROR (include/poutre/base/config.hpp:203-215).
//
// Images live either in host memory (HostImage, what poutre::Create returns, like the
// reference's image_t) or in HBM (DeviceImage).  Operations on device images run entirely
// on the GPU; host images are staged through HBM inside the call.
#pragma once

#include <cstddef>
#include <cstdint>
#include <memory>
#include <string>
#include <variant>
#include <vector>

namespace poutre {

using pUINT8 = std::uint8_t;
using pINT32 = std::int32_t;
using pINT64 = std::int64_t;
using pFLOAT = float;
using pDOUBLE = double;
using pUINT16 = std::uint16_t;

enum class CompoundType { CompoundType_Undef = 0, CompoundType_Scalar = 1 << 0, CompoundType_3Planes = 1 << 1,
                          CompoundType_4Planes = 1 << 2 };

enum class PType {
  PType_Undef = 0,
  PType_GrayUINT8 = 1 << 2,
  PType_GrayINT32 = 1 << 3,
  PType_F32 = 1 << 4,
  PType_GrayINT64 = 1 << 5,
  PType_D64 = 1 << 6,
  PType_GrayUINT16 = 1 << 7,  // extension of this library
};

using ScalarTypeVariant = std::variant<pUINT8, pINT32, pINT64, pFLOAT, pDOUBLE>;

std::size_t SizeOf(PType ptype);

class IInterface
{
public:
  [[nodiscard]] virtual PType GetPType() const = 0;
  [[nodiscard]] virtual CompoundType GetCType() const = 0;
  [[nodiscard]] virtual std::vector<std::size_t> GetShape() const = 0;
  [[nodiscard]] virtual std::size_t GetRank() const = 0;
  virtual ~IInterface() = default;
  [[nodiscard]] virtual std::string str() const = 0;
  // -- access used by this library (no reference counterpart)
  [[nodiscard]] virtual void *Data() const = 0;       // first element
  [[nodiscard]] virtual bool OnDevice() const = 0;    // true if Data() is HBM
};

bool AsSizesCompatible(const IInterface &i_img1, const IInterface &i_img2);
bool AsTypesCompatible(const IInterface &i_img1, const IInterface &i_img2);
bool AreImagesDifferent(const IInterface &i_img1, const IInterface &i_img2);
void AssertSizesCompatible(const IInterface &i_img1, const IInterface &i_img2, const std::string &i_msg);
void AssertAsTypesCompatible(const IInterface &i_img1, const IInterface &i_img2, const std::string &i_msg);
void AssertImagesAreDifferent(const IInterface &i_img1, const IInterface &i_img2, const std::string &i_msg);

// Dense host image (the reference's details::image_t).
class HostImage final : public IInterface
{
public:
  HostImage(std::vector<std::size_t> shape, PType ptype);
  PType GetPType() const override { return m_ptype; }
  CompoundType GetCType() const override { return CompoundType::CompoundType_Scalar; }
  std::vector<std::size_t> GetShape() const override { return m_shape; }
  std::size_t GetRank() const override { return m_shape.size(); }
  std::string str() const override;
  void *Data() const override { return const_cast<unsigned char *>(m_data.data()); }
  bool OnDevice() const override { return false; }
  std::size_t size() const;

private:
  std::vector<std::size_t> m_shape;
  PType m_ptype;
  std::vector<unsigned char> m_data;
};

// Dense image resident in HBM (hipMalloc), owned by the object.
class DeviceImage final : public IInterface
{
public:
  DeviceImage(std::vector<std::size_t> shape, PType ptype);
  ~DeviceImage() override;
  DeviceImage(const DeviceImage &) = delete;
  DeviceImage &operator=(const DeviceImage &) = delete;
  PType GetPType() const override { return m_ptype; }
  CompoundType GetCType() const override { return CompoundType::CompoundType_Scalar; }
  std::vector<std::size_t> GetShape() const override { return m_shape; }
  std::size_t GetRank() const override { return m_shape.size(); }
  std::string str() const override;
  void *Data() const override { return m_data; }
  bool OnDevice() const override { return true; }
  void CopyFrom(const IInterface &src);  // host or device source, same shape/type
  void CopyTo(IInterface &dst) const;

private:
  std::vector<std::size_t> m_shape;
  PType m_ptype;
  void *m_data = nullptr;
};

// Factories: Create gives a host image like the reference; CreateDevice an HBM image.
std::unique_ptr<IInterface> Create(const std::vector<std::size_t> &dims, CompoundType ctype, PType ptype);
std::unique_ptr<IInterface> CreateDevice(const std::vector<std::size_t> &dims, PType ptype);
std::unique_ptr<IInterface> CloneGeometry(const IInterface &i_img);
void CopyInto(const IInterface &i_src, IInterface &o_dst);
// copy_convert.hpp:31-52: Clone / Convert / ConvertGeometry give an image where i_img lives
// (host or device); ConvertInto converts pixel by pixel (a C cast) into o_img's PType
std::unique_ptr<IInterface> Clone(const IInterface &i_img);
std::unique_ptr<IInterface> Convert(const IInterface &i_img, CompoundType ctype, PType ptype);
std::unique_ptr<IInterface> ConvertGeometry(const IInterface &i_img, CompoundType ctype, PType ptype);
std::unique_ptr<IInterface> ConvertGeometry(const IInterface &i_img, PType ptype);
void ConvertInto(const IInterface &i_img, IInterface &o_img);

// KAT fixture format "<Scalar> <GUINT8|GINT32|GINT64|F32|D64> <rank> <dims...> <values...>"
std::unique_ptr<IInterface> ImageFromString(const std::string &str);
std::string ImageToString(const IInterface &i_img);

namespace se {

enum class se_type : std::uint8_t { runtime, neighbor_list, image, chain };

enum class Common_NL_SE : std::uint8_t {
  Undef,
  SESegmentX1D,
  SESquare2D,
  SECross2D,
  SESegmentX2D,
  SESegmentY2D,
  SESegmentX3D,
  SESegmentY3D,
  SESegmentZ3D,
  SECross3D,
  SESquare3D,
};

enum class Compound_NL_SE : std::uint8_t { Undef, Octagon, Rhombicuboctahedron };

class IStructuringElement
{
public:
  virtual ~IStructuringElement() = default;
  [[nodiscard]] virtual std::unique_ptr<IStructuringElement> Transpose() const = 0;
  [[nodiscard]] virtual se_type GetType() const = 0;
  [[nodiscard]] virtual std::unique_ptr<IStructuringElement> RemoveCenter() const = 0;
  [[nodiscard]] virtual size_t GetSize() const = 0;
  [[nodiscard]] virtual std::unique_ptr<IStructuringElement> Clone() const = 0;
  [[nodiscard]] virtual bool is_equal(const IStructuringElement &) const noexcept = 0;
  [[nodiscard]] virtual bool is_equal_unordered(const IStructuringElement &se) const noexcept = 0;
};

// Runtime neighbour list: offsets {dy,dx} / {dz,dy,dx}, centre first when present.
class neighbor_list final : public IStructuringElement
{
public:
  using coord = std::vector<std::ptrdiff_t>;
  neighbor_list(std::size_t rank, std::vector<coord> coords);
  std::unique_ptr<IStructuringElement> Transpose() const override;
  se_type GetType() const override { return se_type::neighbor_list; }
  std::unique_ptr<IStructuringElement> RemoveCenter() const override;
  size_t GetSize() const override { return m_coords.size(); }
  std::unique_ptr<IStructuringElement> Clone() const override;
  bool is_equal(const IStructuringElement &) const noexcept override;
  bool is_equal_unordered(const IStructuringElement &se) const noexcept override;
  // value-returning forms of the reference's neighbor_list_t
  neighbor_list transpose() const;
  neighbor_list remove_center() const;
  bool is_symmetric() const;
  std::pair<neighbor_list, neighbor_list> split_upper_lower() const;
  std::ptrdiff_t maximum_extension(std::size_t dim, bool positive) const;
  std::size_t rank() const { return m_rank; }
  const std::vector<coord> &get_coordinates() const { return m_coords; }

private:
  std::size_t m_rank;
  std::vector<coord> m_coords;
};

// predefined runtime SEs (predefined_nl_se.cpp:19-91)
extern const neighbor_list SESegmentX1D, SESquare2D, SECross2D, SESegmentX2D, SESegmentY2D, SESegmentX3D,
    SESegmentY3D, SESegmentZ3D, SECross3D, SESquare3D;

}  // namespace se

// ---- low-level morphology
void Erode(const IInterface &i_img, se::Common_NL_SE nl_static, const int iter, IInterface &o_img);
void Dilate(const IInterface &i_img, se::Common_NL_SE nl_static, const int iter, IInterface &o_img);
void Erode(const IInterface &i_img, se::Compound_NL_SE nl_compound, const int size, IInterface &o_img);
void Dilate(const IInterface &i_img, se::Compound_NL_SE nl_compound, const int size, IInterface &o_img);
void Erode(const IInterface &i_img, const se::IStructuringElement &str_el, IInterface &o_img);
void Dilate(const IInterface &i_img, const se::IStructuringElement &str_el, IInterface &o_img);
void ErodeX(const IInterface &i_img, const std::ptrdiff_t size_half_segment, IInterface &o_img);
void DilateX(const IInterface &i_img, const std::ptrdiff_t size_half_segment, IInterface &o_img);
void ErodeY(const IInterface &i_img, const std::ptrdiff_t size_half_segment, IInterface &o_img);
void DilateY(const IInterface &i_img, const std::ptrdiff_t size_half_segment, IInterface &o_img);
// extensions: opening / closing / gradient by an iterated Common_NL_SE
void Opening(const IInterface &i_img, se::Common_NL_SE nl_static, const int iter, IInterface &o_img);
void Closing(const IInterface &i_img, se::Common_NL_SE nl_static, const int iter, IInterface &o_img);
void Gradient(const IInterface &i_img, se::Common_NL_SE nl_static, const int iter, IInterface &o_img);

// ---- pixel processing (include/poutre/pixel_processing/arith.hpp:31-53, compare.hpp:31-84).
// Scalars of any ScalarTypeVariant alternative are converted to the images' pixel type (the
// reference requires the alternative of that type).  CompareImage: o_img = (i_img OP comp) ?
// true : false pixel by pixel; comp has i_img's PType, true / false have o_img's, which may
// differ from i_img's.
enum class CompOpType {
  CompOpUndef = 0,
  CompOpEqual = 1 << 0,
  CompOpDiff = 1 << 1,
  CompOpSup = 1 << 2,
  CompOpSupEqual = 1 << 3,
  CompOpInf = 1 << 4,
  CompOpInfEqual = 1 << 5
};
void ArithInvertImage(const IInterface &i_img, IInterface &o_img);
void ArithSupImage(const IInterface &i_img1, const IInterface &i_img2, IInterface &o_img);
void ArithInfImage(const IInterface &i_img1, const IInterface &i_img2, IInterface &o_img);
void ArithSaturatedAddImage(const IInterface &i_img1, const IInterface &i_img2, IInterface &o_img);
void ArithSaturatedSubImage(const IInterface &i_img1, const IInterface &i_img2, IInterface &o_img);
void ArithSaturatedAddConstant(const IInterface &i_img, const ScalarTypeVariant &pvalue, IInterface &o_img);
void ArithSaturatedSubConstant(const IInterface &i_img, const ScalarTypeVariant &pvalue, IInterface &o_img);
void CompareImage(const IInterface &i_img, CompOpType compOpType, const ScalarTypeVariant &i_comp,
                  const ScalarTypeVariant &i_valtrue, const ScalarTypeVariant &i_valfalse, IInterface &o_img);
void CompareImage(const IInterface &i_img, CompOpType compOpType, const IInterface &i_imgcomp,
                  const IInterface &i_imgtrue, const IInterface &i_imgfalse, IInterface &o_img);
void CompareImage(const IInterface &i_img, CompOpType compOpType, const ScalarTypeVariant &i_comp,
                  const IInterface &i_imgtrue, const IInterface &i_imgfalse, IInterface &o_img);
void CompareImage(const IInterface &i_img, CompOpType compOpType, const ScalarTypeVariant &i_comp,
                  const IInterface &i_imgtrue, const ScalarTypeVariant &i_valfalse, IInterface &o_img);
void CompareImage(const IInterface &i_img, CompOpType compOpType, const IInterface &i_imgcomp,
                  const ScalarTypeVariant &i_valtrue, const IInterface &i_imgfalse, IInterface &o_img);
void CompareImage(const IInterface &i_img, CompOpType compOpType, const IInterface &i_imgcomp,
                  const ScalarTypeVariant &i_valtrue, const ScalarTypeVariant &i_valfalse, IInterface &o_img);
void CompareImage(const IInterface &i_img, CompOpType compOpType, const ScalarTypeVariant &i_comp,
                  const ScalarTypeVariant &i_valtrue, const IInterface &i_imgfalse, IInterface &o_img);
void CompareImage(const IInterface &i_img, CompOpType compOpType, const IInterface &i_imgcomp,
                  const IInterface &i_imgtrue, const ScalarTypeVariant &i_valfalse, IInterface &o_img);

namespace geo {

enum class reconstruction_type : std::uint8_t { erode, dilate };

void Reconstruction(reconstruction_type rect_type, const IInterface &i_marker, const IInterface &i_mask,
                    se::Common_NL_SE nl_static, IInterface &o_img);
void low_leveling(const IInterface &i_ref, const IInterface &i_marker, se::Common_NL_SE nl_static, IInterface &o_img);
void high_leveling(const IInterface &i_ref, const IInterface &i_marker, se::Common_NL_SE nl_static, IInterface &o_img);
void leveling(const IInterface &i_ref, const IInterface &i_marker, se::Common_NL_SE nl_static, IInterface &o_img);
void h_minima(const IInterface &i_img, const ScalarTypeVariant &pvalue, se::Common_NL_SE nl_static, IInterface &o_img);
void h_maxima(const IInterface &i_img, const ScalarTypeVariant &pvalue, se::Common_NL_SE nl_static, IInterface &o_img);
void h_convex(const IInterface &i_img, const ScalarTypeVariant &pvalue, se::Common_NL_SE nl_static, IInterface &o_img);
void h_concave(const IInterface &i_img, const ScalarTypeVariant &pvalue, se::Common_NL_SE nl_static, IInterface &o_img);
void dynamic_pseudo_opening(const IInterface &i_img, const ScalarTypeVariant &pvalue, se::Common_NL_SE nl_static,
                            IInterface &o_img);
void dynamic_pseudo_closing(const IInterface &i_img, const ScalarTypeVariant &pvalue, se::Common_NL_SE nl_static,
                            IInterface &o_img);

}  // namespace geo

namespace label {
// include/poutre/label/label_binary.hpp:35, label_flat_zones.hpp:35, label_extrema.hpp:34-39;
// o_img must be a GINT64 image of i_img's shape
std::size_t label_binary(const IInterface &i_img, se::Common_NL_SE nl_static, IInterface &o_img);
std::size_t label_flat_zones(const IInterface &i_img, se::Common_NL_SE nl_static, IInterface &o_img);
void label_minima(const IInterface &i_img, se::Common_NL_SE nl_static, IInterface &o_img);
void label_maxima(const IInterface &i_img, se::Common_NL_SE nl_static, IInterface &o_img);
}  // namespace label
}  // namespace poutre
