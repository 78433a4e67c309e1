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
//   details::av::index<R> (idx1d..idx3d)    include/poutre/base/details/data_structures/array_view.hpp:335-485
//   se::details::neighbor_list_t<R> + SESquare2D ... predefined objects
//                                           neighbor_list_se_t.hpp:36-228, predefined_nl_se.hpp:41-99,
//                                           predefined_nl_se.cpp:19-91
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

#include <algorithm>
#include <array>
#include <cstddef>
#include <cstdint>
#include <initializer_list>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
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

// Dense host image (the reference's details::image_t), rank 1..4.  Compound images
// (3Planes / 4Planes, the reference's compound_type<T, N>) store the N planes of a pixel
// next to each other; the morphology and geodesy operators take Scalar images only.
class HostImage final : public IInterface
{
public:
  HostImage(std::vector<std::size_t> shape, PType ptype,
            CompoundType ctype = CompoundType::CompoundType_Scalar);
  PType GetPType() const override { return m_ptype; }
  CompoundType GetCType() const override { return m_ctype; }
  std::vector<std::size_t> GetShape() const override { return m_shape; }
  std::size_t GetRank() const override { return m_shape.size(); }
  std::string str() const override;
  void *Data() const override { return const_cast<unsigned char *>(m_data.data()); }
  bool OnDevice() const override { return false; }
  std::size_t size() const;    // pixels
  std::size_t planes() const;  // values per pixel: 1, 3 or 4

private:
  std::vector<std::size_t> m_shape;
  PType m_ptype;
  CompoundType m_ctype;
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

// KAT fixture format "<Scalar|3Planes|4Planes> <GUINT8|GINT32|GINT64|F32|D64> <rank> <dims...>
// <values...>" (image_interface.cpp:616-909): each value is read and written through its
// PType's text type (unsigned int for GUINT8, int, int64, float, double), values of a
// compound pixel consecutive, separated by single spaces, default stream formatting.
std::unique_ptr<IInterface> ImageFromString(const std::string &str);
std::string ImageToString(const IInterface &i_img);

namespace details::av {

// Coordinates of rank Rank (array_view.hpp:335-485).  operator< orders from the LAST
// coordinate to the first; operator<= / >= compare every coordinate (a partial order), as in
// the reference.
template <std::ptrdiff_t Rank> class index
{
public:
  static constexpr std::ptrdiff_t rank = Rank;
  using value_type = std::ptrdiff_t;
  using size_type = std::size_t;
  using reference = value_type &;
  using const_reference = const value_type &;

  constexpr index() noexcept = default;
  constexpr index(value_type v) noexcept { m_c.fill(v); }  // NOLINT: every coordinate = v
  constexpr index(std::initializer_list<value_type> il)
  {
    if (il.size() != static_cast<std::size_t>(Rank)) throw std::runtime_error("Throw index: wrong number of coordinates");
    std::copy(il.begin(), il.end(), m_c.begin());
  }
  constexpr reference operator[](std::size_t n) { return m_c[n]; }
  constexpr const_reference operator[](std::size_t n) const { return m_c[n]; }
  constexpr index &operator+=(const index &rhs)
  {
    for (std::size_t i = 0; i < static_cast<std::size_t>(Rank); i++) m_c[i] += rhs.m_c[i];
    return *this;
  }
  constexpr index &operator-=(const index &rhs)
  {
    for (std::size_t i = 0; i < static_cast<std::size_t>(Rank); i++) m_c[i] -= rhs.m_c[i];
    return *this;
  }
  constexpr index &operator*=(value_type v)
  {
    for (auto &c : m_c) c *= v;
    return *this;
  }
  constexpr index &operator/=(value_type v)
  {
    for (auto &c : m_c) c /= v;
    return *this;
  }
  constexpr index operator+() const noexcept { return *this; }
  constexpr index operator-() const
  {
    index r = *this;
    for (auto &c : r.m_c) c = -c;
    return r;
  }
  constexpr bool operator<(const index &rhs) const noexcept
  {
    for (std::ptrdiff_t i = Rank - 1; i >= 0; --i)
      if (m_c[(std::size_t)i] != rhs.m_c[(std::size_t)i]) return m_c[(std::size_t)i] < rhs.m_c[(std::size_t)i];
    return false;
  }
  constexpr bool operator>(const index &rhs) const noexcept { return rhs < *this; }
  constexpr bool operator<=(const index &rhs) const noexcept
  {
    for (std::size_t i = 0; i < static_cast<std::size_t>(Rank); i++)
      if (m_c[i] > rhs.m_c[i]) return false;
    return true;
  }
  constexpr bool operator>=(const index &rhs) const noexcept
  {
    for (std::size_t i = 0; i < static_cast<std::size_t>(Rank); i++)
      if (m_c[i] < rhs.m_c[i]) return false;
    return true;
  }
  friend constexpr bool operator==(const index &a, const index &b) noexcept { return a.m_c == b.m_c; }
  friend constexpr bool operator!=(const index &a, const index &b) noexcept { return !(a == b); }

private:
  std::array<value_type, static_cast<std::size_t>(Rank)> m_c = {};
};
using idx1d = index<1>;
using idx2d = index<2>;
using idx3d = index<3>;
using idx4d = index<4>;

}  // namespace details::av

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

namespace details {

// Runtime neighbour list (neighbor_list_se_t.hpp:36-228): offsets {dy, dx} / {dz, dy, dx}, the
// centre first.  The reference takes the centre of split_upper_lower / transpose /
// remove_center from a function-local static initialised by the first call of any list of
// that rank; here it is the first offset of the list being used (the two agree for every
// list whose first offset is the origin, as for all predefined SEs).
template <std::ptrdiff_t Rank> class neighbor_list_t final : public IStructuringElement
{
public:
  static const std::ptrdiff_t rank = Rank;
  using neighbor_element = poutre::details::av::index<Rank>;
  using self_type = neighbor_list_t<Rank>;
  using storage_type = std::vector<neighbor_element>;
  using const_iterator = typename storage_type::const_iterator;

  explicit neighbor_list_t(const storage_type &v) : m_c(v) {}
  neighbor_list_t(std::initializer_list<neighbor_element> il) : m_c(il) {}

  const_iterator begin() const { return m_c.begin(); }
  const_iterator end() const { return m_c.end(); }
  const storage_type &get_coordinates() const noexcept { return m_c; }
  std::size_t size() const { return m_c.size(); }

  bool is_symmetric() const noexcept
  {
    auto s = m_c;
    std::sort(s.begin(), s.end());
    for (std::size_t i = 0; i < s.size() / 2; ++i)
      if (s[i] != -s[s.size() - i - 1]) return false;
    return true;
  }
  // (upper, lower): the offsets before / after the centre in the order of operator<
  std::pair<self_type, self_type> split_upper_lower() const
  {
    if (!is_symmetric()) throw std::runtime_error("Throw split_upper_lower the SE is not symmetric");
    const neighbor_element c = centre();
    storage_type up, lo;
    for (const auto &e : m_c) {
      if (e == c) continue;
      if (e < c) up.push_back(e);
      else lo.push_back(e);
    }
    return {self_type(up), self_type(lo)};
  }
  self_type transpose() const
  {
    const neighbor_element c = centre();
    storage_type out(m_c.size());
    for (std::size_t i = 0; i < m_c.size(); ++i)
      for (std::size_t j = 0; j < static_cast<std::size_t>(Rank); ++j) out[i][j] = 2 * c[j] - m_c[i][j];
    return self_type(out);
  }
  self_type remove_center() const
  {
    const neighbor_element c = centre();
    storage_type out;
    for (const auto &e : m_c)
      if (e != c) out.push_back(e);
    return self_type(out);
  }
  // largest extension along `dimension`, forward (true) or backward; 0 at least
  std::ptrdiff_t maximum_extension(const std::size_t dimension, const bool direction) const
  {
    std::ptrdiff_t ext = 0;
    for (const auto &e : m_c) ext = std::max(ext, direction ? e[dimension] : -e[dimension]);
    return ext;
  }
  // (smallest, largest) coordinate along every dimension
  std::pair<neighbor_element, neighbor_element> maximum_extension() const
  {
    neighbor_element mn(std::numeric_limits<std::ptrdiff_t>::max()), mx(std::numeric_limits<std::ptrdiff_t>::min());
    for (const auto &e : m_c)
      for (std::size_t d = 0; d < static_cast<std::size_t>(Rank); d++) {
        mn[d] = std::min(mn[d], e[d]);
        mx[d] = std::max(mx[d], e[d]);
      }
    return {mn, mx};
  }
  bool operator==(const self_type &rhs) const { return (this == &rhs) || (m_c == rhs.m_c); }
  bool operator!=(const self_type &rhs) const { return !(*this == rhs); }
  // the same SETS of offsets (sorted, repeats removed)
  bool is_equal_unordered(const self_type &rhs) const
  {
    auto a = m_c, b = rhs.m_c;
    std::sort(a.begin(), a.end());
    std::sort(b.begin(), b.end());
    a.erase(std::unique(a.begin(), a.end()), a.end());
    b.erase(std::unique(b.begin(), b.end()), b.end());
    return a == b;
  }

  // IStructuringElement
  se_type GetType() const override { return se_type::neighbor_list; }
  size_t GetSize() const override { return size(); }
  std::unique_ptr<IStructuringElement> Transpose() const override { return std::make_unique<self_type>(transpose()); }
  std::unique_ptr<IStructuringElement> RemoveCenter() const override
  { return std::make_unique<self_type>(remove_center()); }
  std::unique_ptr<IStructuringElement> Clone() const override { return std::make_unique<self_type>(*this); }
  bool is_equal(const IStructuringElement &se) const noexcept override
  {
    const auto *p = dynamic_cast<const self_type *>(&se);
    return p != nullptr && *this == *p;
  }
  bool is_equal_unordered(const IStructuringElement &se) const noexcept override
  {
    const auto *p = dynamic_cast<const self_type *>(&se);
    return p != nullptr && is_equal_unordered(*p);
  }

private:
  neighbor_element centre() const { return m_c.empty() ? neighbor_element(0) : m_c[0]; }
  storage_type m_c;
};

}  // namespace details

// predefined runtime SEs (predefined_nl_se.hpp:41-99; offsets in the order of
// predefined_nl_se.cpp:19-91)
extern const details::neighbor_list_t<1> SESegmentX1D;
extern const details::neighbor_list_t<2> SESquare2D, SECross2D, SESegmentX2D, SESegmentY2D;
extern const details::neighbor_list_t<3> SESegmentX3D, SESegmentY3D, SESegmentZ3D, SECross3D, SESquare3D;

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
