// poutre_api.cpp -- C++ API of include/poutre_amd.hpp on top of the C ABI (include/poutre_amd.h).
//
// Each function performs the reference's argument checks on the IInterface objects
// (sizes, types, distinct objects; include/poutre/base/image_interface.hpp:95-102),
// stages host images through HBM when needed, calls the pm_* entry point on the default
// stream and converts a negative status into std::runtime_error("Throw <message>").
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <sstream>
#include <stdexcept>

#include "../../include/poutre_amd.h"
#include "../../include/poutre_amd.hpp"

namespace poutre {

[[noreturn]] static void fail(const std::string &msg) { throw std::runtime_error("Throw " + msg); }

static void hip_check(hipError_t e, const char *what)
{
  if (e != hipSuccess) fail(std::string(what) + ": " + hipGetErrorString(e));
}

static void pm_check(int rc)
{
  if (rc != PM_OK) fail(pm_last_error());
}

std::size_t SizeOf(PType p)
{
  switch (p) {
  case PType::PType_GrayUINT8: return 1;
  case PType::PType_GrayUINT16: return 2;
  case PType::PType_GrayINT32: return 4;
  case PType::PType_F32: return 4;
  case PType::PType_GrayINT64: return 8;
  case PType::PType_D64: return 8;
  default: fail("Unsupported PType");
  }
}

static std::size_t count(const std::vector<std::size_t> &s)
{
  std::size_t n = 1;
  for (auto d : s) n *= d;
  return n;
}

// ------------------------------------------------------------------ checks
bool AsSizesCompatible(const IInterface &a, const IInterface &b) { return a.GetShape() == b.GetShape(); }
bool AsTypesCompatible(const IInterface &a, const IInterface &b)
{
  return a.GetPType() == b.GetPType() && a.GetCType() == b.GetCType();
}
bool AreImagesDifferent(const IInterface &a, const IInterface &b) { return &a != &b; }
void AssertSizesCompatible(const IInterface &a, const IInterface &b, const std::string &m)
{
  if (!AsSizesCompatible(a, b)) fail(m);
}
void AssertAsTypesCompatible(const IInterface &a, const IInterface &b, const std::string &m)
{
  if (!AsTypesCompatible(a, b)) fail(m);
}
void AssertImagesAreDifferent(const IInterface &a, const IInterface &b, const std::string &m)
{
  if (!AreImagesDifferent(a, b)) fail(m);
}

// ------------------------------------------------------------------ images
static const char *ptype_name(PType p)
{
  switch (p) {
  case PType::PType_GrayUINT8: return "GUINT8";
  case PType::PType_GrayUINT16: return "GUINT16";
  case PType::PType_GrayINT32: return "GINT32";
  case PType::PType_GrayINT64: return "GINT64";
  case PType::PType_F32: return "F32";
  case PType::PType_D64: return "D64";
  default: return "Undef";
  }
}

static std::string describe(const char *kind, const std::vector<std::size_t> &s, PType p)
{
  std::ostringstream os;
  os << kind << " Scalar " << ptype_name(p) << " rank " << s.size() << " shape";
  for (auto d : s) os << ' ' << d;
  return os.str();
}

HostImage::HostImage(std::vector<std::size_t> shape, PType ptype)
    : m_shape(std::move(shape)), m_ptype(ptype), m_data(count(m_shape) * SizeOf(ptype))
{
  if (m_shape.empty() || m_shape.size() > 3) fail("Unsupported number of dims");
}
std::string HostImage::str() const { return describe("HostImage", m_shape, m_ptype); }
std::size_t HostImage::size() const { return count(m_shape); }

DeviceImage::DeviceImage(std::vector<std::size_t> shape, PType ptype) : m_shape(std::move(shape)), m_ptype(ptype)
{
  if (m_shape.empty() || m_shape.size() > 3) fail("Unsupported number of dims");
  const std::size_t bytes = std::max<std::size_t>(count(m_shape) * SizeOf(ptype), 1);
  hip_check(hipMalloc(&m_data, bytes), "hipMalloc");
}
DeviceImage::~DeviceImage()
{
  if (m_data) (void)hipFree(m_data);
}
std::string DeviceImage::str() const { return describe("DeviceImage", m_shape, m_ptype); }

static void copy_any(const IInterface &src, const IInterface &dst)
{
  AssertSizesCompatible(src, dst, "CopyInto images have not compatible sizes");
  AssertAsTypesCompatible(src, dst, "CopyInto images must have compatible types");
  const std::size_t bytes = count(src.GetShape()) * SizeOf(src.GetPType());
  if (bytes == 0) return;
  hipMemcpyKind kind = src.OnDevice() ? (dst.OnDevice() ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost)
                                      : (dst.OnDevice() ? hipMemcpyHostToDevice : hipMemcpyHostToHost);
  hip_check(hipMemcpy(dst.Data(), src.Data(), bytes, kind), "hipMemcpy");
}
void DeviceImage::CopyFrom(const IInterface &src) { copy_any(src, *this); }
void DeviceImage::CopyTo(IInterface &dst) const { copy_any(*this, dst); }
void CopyInto(const IInterface &i_src, IInterface &o_dst) { copy_any(i_src, o_dst); }

std::unique_ptr<IInterface> Create(const std::vector<std::size_t> &dims, CompoundType ctype, PType ptype)
{
  if (ctype != CompoundType::CompoundType_Scalar) fail("Create: only scalar images are supported");
  return std::make_unique<HostImage>(dims, ptype);
}
std::unique_ptr<IInterface> CreateDevice(const std::vector<std::size_t> &dims, PType ptype)
{
  return std::make_unique<DeviceImage>(dims, ptype);
}
std::unique_ptr<IInterface> CloneGeometry(const IInterface &i_img)
{
  if (i_img.OnDevice()) return CreateDevice(i_img.GetShape(), i_img.GetPType());
  return Create(i_img.GetShape(), i_img.GetCType(), i_img.GetPType());
}

// ------------------------------------------------------------------ fixture format
static PType ptype_from(const std::string &s)
{
  if (s == "GUINT8") return PType::PType_GrayUINT8;
  if (s == "GUINT16") return PType::PType_GrayUINT16;
  if (s == "GINT32") return PType::PType_GrayINT32;
  if (s == "GINT64") return PType::PType_GrayINT64;
  if (s == "F32") return PType::PType_F32;
  if (s == "D64") return PType::PType_D64;
  fail("ImageFromString: unknown PType " + s);
}

template <typename T> static void fill_from(std::istringstream &is, void *data, std::size_t n)
{
  T *p = static_cast<T *>(data);
  for (std::size_t i = 0; i < n; i++) {
    double v;
    if (!(is >> v)) fail("ImageFromString: not enough values");
    p[i] = static_cast<T>(v);
  }
}

std::unique_ptr<IInterface> ImageFromString(const std::string &str)
{
  std::istringstream is(str);
  std::string ctype, pt;
  std::size_t rank = 0;
  is >> ctype >> pt >> rank;
  if (ctype != "Scalar") fail("ImageFromString: only Scalar images are supported");
  std::vector<std::size_t> shape(rank);
  for (auto &d : shape) is >> d;
  const PType p = ptype_from(pt);
  auto img = Create(shape, CompoundType::CompoundType_Scalar, p);
  const std::size_t n = count(shape);
  switch (p) {
  case PType::PType_GrayUINT8: fill_from<pUINT8>(is, img->Data(), n); break;
  case PType::PType_GrayUINT16: fill_from<pUINT16>(is, img->Data(), n); break;
  case PType::PType_GrayINT32: fill_from<pINT32>(is, img->Data(), n); break;
  case PType::PType_GrayINT64: fill_from<pINT64>(is, img->Data(), n); break;
  case PType::PType_F32: fill_from<pFLOAT>(is, img->Data(), n); break;
  case PType::PType_D64: fill_from<pDOUBLE>(is, img->Data(), n); break;
  default: break;
  }
  return img;  // trailing values, if any, are ignored as by the reference's parser
}

template <typename T> static void print_to(std::ostringstream &os, const void *data, std::size_t n)
{
  const T *p = static_cast<const T *>(data);
  for (std::size_t i = 0; i < n; i++) {
    os << ' ';
    if constexpr (std::is_floating_point<T>::value) os << p[i];
    else os << static_cast<long long>(p[i]);
  }
}

std::string ImageToString(const IInterface &i_img)
{
  std::unique_ptr<IInterface> host;
  const IInterface *src = &i_img;
  if (i_img.OnDevice()) {
    host = Create(i_img.GetShape(), CompoundType::CompoundType_Scalar, i_img.GetPType());
    CopyInto(i_img, *host);
    src = host.get();
  }
  std::ostringstream os;
  os << "Scalar " << ptype_name(src->GetPType()) << ' ' << src->GetRank();
  for (auto d : src->GetShape()) os << ' ' << d;
  const std::size_t n = count(src->GetShape());
  switch (src->GetPType()) {
  case PType::PType_GrayUINT8: print_to<pUINT8>(os, src->Data(), n); break;
  case PType::PType_GrayUINT16: print_to<pUINT16>(os, src->Data(), n); break;
  case PType::PType_GrayINT32: print_to<pINT32>(os, src->Data(), n); break;
  case PType::PType_GrayINT64: print_to<pINT64>(os, src->Data(), n); break;
  case PType::PType_F32: print_to<pFLOAT>(os, src->Data(), n); break;
  case PType::PType_D64: print_to<pDOUBLE>(os, src->Data(), n); break;
  default: break;
  }
  return os.str();
}

// ------------------------------------------------------------------ structuring elements
namespace se {

neighbor_list::neighbor_list(std::size_t rank, std::vector<coord> coords) : m_rank(rank), m_coords(std::move(coords))
{
  for (const auto &c : m_coords)
    if (c.size() != rank) fail("neighbor_list: coordinate of wrong rank");
}
neighbor_list neighbor_list::transpose() const
{
  std::vector<coord> t = m_coords;
  for (auto &c : t)
    for (auto &v : c) v = -v;
  return neighbor_list(m_rank, t);
}
neighbor_list neighbor_list::remove_center() const
{
  std::vector<coord> t;
  for (const auto &c : m_coords)
    if (std::any_of(c.begin(), c.end(), [](std::ptrdiff_t v) { return v != 0; })) t.push_back(c);
  return neighbor_list(m_rank, t);
}
bool neighbor_list::is_symmetric() const { return transpose().is_equal_unordered(*this); }
std::pair<neighbor_list, neighbor_list> neighbor_list::split_upper_lower() const
{
  std::vector<coord> up, lo;
  for (const auto &c : m_coords) {
    int s = 0;
    for (auto v : c)
      if (v != 0) { s = v < 0 ? -1 : 1; break; }
    if (s < 0) up.push_back(c);
    else if (s > 0) lo.push_back(c);
  }
  return {neighbor_list(m_rank, up), neighbor_list(m_rank, lo)};
}
std::ptrdiff_t neighbor_list::maximum_extension(std::size_t dim, bool positive) const
{
  std::ptrdiff_t e = 0;
  for (const auto &c : m_coords) e = std::max(e, positive ? c.at(dim) : -c.at(dim));
  return e;
}
std::unique_ptr<IStructuringElement> neighbor_list::Transpose() const
{ return std::make_unique<neighbor_list>(transpose()); }
std::unique_ptr<IStructuringElement> neighbor_list::RemoveCenter() const
{ return std::make_unique<neighbor_list>(remove_center()); }
std::unique_ptr<IStructuringElement> neighbor_list::Clone() const { return std::make_unique<neighbor_list>(*this); }
bool neighbor_list::is_equal(const IStructuringElement &o) const noexcept
{
  auto p = dynamic_cast<const neighbor_list *>(&o);
  return p && p->m_rank == m_rank && p->m_coords == m_coords;
}
// the same SETS of offsets (neighbor_list_se_t.hpp:177-195: sorted, repeats removed)
bool neighbor_list::is_equal_unordered(const IStructuringElement &o) const noexcept
{
  auto p = dynamic_cast<const neighbor_list *>(&o);
  if (!p || p->m_rank != m_rank) return false;
  auto a = m_coords, b = p->m_coords;
  std::sort(a.begin(), a.end());
  std::sort(b.begin(), b.end());
  a.erase(std::unique(a.begin(), a.end()), a.end());
  b.erase(std::unique(b.begin(), b.end()), b.end());
  return a == b;
}

static neighbor_list make_box(std::size_t rank, bool cross)
{
  std::vector<neighbor_list::coord> c;
  c.push_back(neighbor_list::coord(rank, 0));
  const int n = rank == 2 ? 9 : 27;
  for (int i = 0; i < n; i++) {
    neighbor_list::coord p(rank);
    int k = i, l1 = 0;
    for (std::size_t a = 0; a < rank; a++) {
      p[rank - 1 - a] = (k % 3) - 1;
      l1 += std::abs(p[rank - 1 - a]);
      k /= 3;
    }
    if (l1 == 0 || (cross && l1 > 1)) continue;
    c.push_back(p);
  }
  return neighbor_list(rank, c);
}
static neighbor_list make_seg(std::size_t rank, std::size_t axis)
{
  std::vector<neighbor_list::coord> c(3, neighbor_list::coord(rank, 0));
  c[1][axis] = -1;
  c[2][axis] = 1;
  return neighbor_list(rank, c);
}
const neighbor_list SESegmentX1D = make_seg(1, 0);
const neighbor_list SESquare2D = make_box(2, false);
const neighbor_list SECross2D = make_box(2, true);
const neighbor_list SESegmentX2D = make_seg(2, 1);
const neighbor_list SESegmentY2D = make_seg(2, 0);
const neighbor_list SESegmentX3D = make_seg(3, 2);
const neighbor_list SESegmentY3D = make_seg(3, 1);
const neighbor_list SESegmentZ3D = make_seg(3, 0);
const neighbor_list SECross3D = make_box(3, true);
const neighbor_list SESquare3D = make_box(3, false);

}  // namespace se

// ------------------------------------------------------------------ staging
namespace {

// Device view of an image: the image itself when it lives in HBM, a staged copy otherwise.
class Staged
{
public:
  Staged(const IInterface &img, bool upload) : m_img(img)
  {
    m_desc.ptype = (int32_t)img.GetPType();
    m_desc.rank = (int32_t)img.GetRank();
    const auto s = img.GetShape();
    for (std::size_t i = 0; i < 3; i++) m_desc.shape[i] = i < s.size() ? (int64_t)s[i] : 0;
    if (img.OnDevice()) {
      m_desc.data = img.Data();
      return;
    }
    m_bytes = count(s) * SizeOf(img.GetPType());
    hip_check(hipMalloc(&m_tmp, std::max<std::size_t>(m_bytes, 1)), "hipMalloc");
    m_desc.data = m_tmp;
    if (upload && m_bytes) hip_check(hipMemcpy(m_tmp, img.Data(), m_bytes, hipMemcpyHostToDevice), "hipMemcpy");
  }
  ~Staged()
  {
    if (m_tmp) (void)hipFree(m_tmp);
  }
  pm_image *desc() { return &m_desc; }
  void download()
  {
    if (m_tmp && m_bytes) hip_check(hipMemcpy(m_img.Data(), m_tmp, m_bytes, hipMemcpyDeviceToHost), "hipMemcpy");
  }

private:
  const IInterface &m_img;
  pm_image m_desc{};
  void *m_tmp = nullptr;
  std::size_t m_bytes = 0;
};

void checks(const IInterface &a, const IInterface &out, const std::string &what)
{
  AssertSizesCompatible(a, out, what + " images have not compatible sizes");
  AssertAsTypesCompatible(a, out, what + " images must have compatible types");
  AssertImagesAreDifferent(a, out, what + " input ouput images must be different");
}

template <typename F> void unary(const IInterface &in, IInterface &out, const std::string &what, F &&f)
{
  checks(in, out, what);
  Staged si(in, true), so(out, false);
  pm_check(f(si.desc(), so.desc()));
  so.download();
}

template <typename F>
void binary(const IInterface &a, const IInterface &b, IInterface &out, const std::string &what, F &&f)
{
  checks(a, out, what);
  checks(b, out, what);
  AssertSizesCompatible(a, b, what + " images have not compatible sizes");
  AssertAsTypesCompatible(a, b, what + " images must have compatible types");
  Staged sa(a, true), sb(b, true), so(out, false);
  pm_check(f(sa.desc(), sb.desc(), so.desc()));
  so.download();
}

double scalar(const ScalarTypeVariant &v)
{
  return std::visit([](auto x) { return static_cast<double>(x); }, v);
}

}  // namespace

// ------------------------------------------------------------------ pixel processing
// The reference's arith and compare functions check sizes and types only (arith.cpp,
// compare.cpp): the output may be an input.
namespace {
template <typename F>
void arith_binary(const IInterface &a, const IInterface &b, IInterface &out, const std::string &what, F &&f)
{
  AssertSizesCompatible(a, out, what + " images have not compatible sizes");
  AssertSizesCompatible(b, out, what + " images have not compatible sizes");
  AssertAsTypesCompatible(a, out, what + " images must have compatible types");
  AssertAsTypesCompatible(b, out, what + " images must have compatible types");
  Staged sa(a, true), sb(b, true), so(out, false);
  pm_check(f(sa.desc(), sb.desc(), so.desc()));
  so.download();
}
template <typename F> void arith_unary(const IInterface &a, IInterface &out, const std::string &what, F &&f)
{
  AssertSizesCompatible(a, out, what + " images have not compatible sizes");
  AssertAsTypesCompatible(a, out, what + " images must have compatible types");
  Staged sa(a, true), so(out, false);
  pm_check(f(sa.desc(), so.desc()));
  so.download();
}

// an operand of CompareImage: an image or a scalar
struct CmpArg {
  const IInterface *img = nullptr;
  double value = 0;
  CmpArg(const IInterface &i) : img(&i) {}
  CmpArg(const ScalarTypeVariant &v) : value(scalar(v)) {}
};

void compare_any(const IInterface &in, CompOpType op, const CmpArg &c, const CmpArg &t, const CmpArg &f,
                 IInterface &out)
{
  AssertSizesCompatible(in, out, "CompareImage images have not compatible sizes");
  for (const CmpArg *x : {&c, &t, &f})
    if (x->img) AssertSizesCompatible(*x->img, out, "CompareImage images have not compatible sizes");
  Staged si(in, true), so(out, false);
  std::unique_ptr<Staged> st[3];
  pm_operand po[3];
  const CmpArg *args[3] = {&c, &t, &f};
  for (int k = 0; k < 3; k++) {
    po[k].image = nullptr;
    po[k].value = args[k]->value;
    if (args[k]->img) {
      st[k] = std::make_unique<Staged>(*args[k]->img, true);
      po[k].image = st[k]->desc();
    }
  }
  pm_check(pm_compare(si.desc(), (int)op, &po[0], &po[1], &po[2], so.desc(), nullptr));
  so.download();
}
}  // namespace

std::unique_ptr<IInterface> ConvertGeometry(const IInterface &i, CompoundType ctype, PType ptype)
{
  if (ctype != CompoundType::CompoundType_Scalar) fail("ConvertGeometry: only scalar images are supported");
  if (i.OnDevice()) return CreateDevice(i.GetShape(), ptype);
  return Create(i.GetShape(), ctype, ptype);
}
std::unique_ptr<IInterface> ConvertGeometry(const IInterface &i, PType ptype)
{
  return ConvertGeometry(i, CompoundType::CompoundType_Scalar, ptype);
}
void ConvertInto(const IInterface &i, IInterface &o)
{
  AssertSizesCompatible(i, o, "ConvertInto images have not compatible sizes");
  Staged si(i, true), so(o, false);
  pm_check(pm_convert_into(si.desc(), so.desc(), nullptr));
  so.download();
}
std::unique_ptr<IInterface> Convert(const IInterface &i, CompoundType ctype, PType ptype)
{
  auto o = ConvertGeometry(i, ctype, ptype);
  ConvertInto(i, *o);
  return o;
}
std::unique_ptr<IInterface> Clone(const IInterface &i)
{
  auto o = CloneGeometry(i);
  CopyInto(i, *o);
  return o;
}

void ArithInvertImage(const IInterface &i, IInterface &o)
{
  arith_unary(i, o, "ArithInvertImage", [&](pm_image *a, pm_image *b) { return pm_arith_invert(a, b, nullptr); });
}
void ArithSupImage(const IInterface &a, const IInterface &b, IInterface &o)
{
  arith_binary(a, b, o, "ArithSupImage", [&](pm_image *x, pm_image *y, pm_image *z) { return pm_arith_sup(x, y, z, nullptr); });
}
void ArithInfImage(const IInterface &a, const IInterface &b, IInterface &o)
{
  arith_binary(a, b, o, "ArithInfImage", [&](pm_image *x, pm_image *y, pm_image *z) { return pm_arith_inf(x, y, z, nullptr); });
}
void ArithSaturatedAddImage(const IInterface &a, const IInterface &b, IInterface &o)
{
  arith_binary(a, b, o, "ArithSaturatedAddImage",
               [&](pm_image *x, pm_image *y, pm_image *z) { return pm_arith_sadd_image(x, y, z, nullptr); });
}
void ArithSaturatedSubImage(const IInterface &a, const IInterface &b, IInterface &o)
{
  arith_binary(a, b, o, "ArithSaturatedSubImage",
               [&](pm_image *x, pm_image *y, pm_image *z) { return pm_arith_ssub_image(x, y, z, nullptr); });
}
void ArithSaturatedAddConstant(const IInterface &i, const ScalarTypeVariant &v, IInterface &o)
{
  arith_unary(i, o, "ArithSaturatedAddConstant",
              [&](pm_image *a, pm_image *b) { return pm_arith_sadd_const(a, scalar(v), b, nullptr); });
}
void ArithSaturatedSubConstant(const IInterface &i, const ScalarTypeVariant &v, IInterface &o)
{
  arith_unary(i, o, "ArithSaturatedSubConstant",
              [&](pm_image *a, pm_image *b) { return pm_arith_ssub_const(a, scalar(v), b, nullptr); });
}
#define PM_CMP_OVERLOAD(TC, TT, TF)                                                                           \
  void CompareImage(const IInterface &i, CompOpType op, TC c, TT t, TF f, IInterface &o)                       \
  {                                                                                                           \
    compare_any(i, op, CmpArg(c), CmpArg(t), CmpArg(f), o);                                                   \
  }
PM_CMP_OVERLOAD(const ScalarTypeVariant &, const ScalarTypeVariant &, const ScalarTypeVariant &)
PM_CMP_OVERLOAD(const IInterface &, const IInterface &, const IInterface &)
PM_CMP_OVERLOAD(const ScalarTypeVariant &, const IInterface &, const IInterface &)
PM_CMP_OVERLOAD(const ScalarTypeVariant &, const IInterface &, const ScalarTypeVariant &)
PM_CMP_OVERLOAD(const IInterface &, const ScalarTypeVariant &, const IInterface &)
PM_CMP_OVERLOAD(const IInterface &, const ScalarTypeVariant &, const ScalarTypeVariant &)
PM_CMP_OVERLOAD(const ScalarTypeVariant &, const ScalarTypeVariant &, const IInterface &)
PM_CMP_OVERLOAD(const IInterface &, const IInterface &, const ScalarTypeVariant &)
#undef PM_CMP_OVERLOAD

// ------------------------------------------------------------------ low-level morphology
void Erode(const IInterface &i, se::Common_NL_SE s, const int iter, IInterface &o)
{
  unary(i, o, "Erode iter", [&](pm_image *a, pm_image *b) { return pm_erode_static(a, (int)s, iter, b, nullptr); });
}
void Dilate(const IInterface &i, se::Common_NL_SE s, const int iter, IInterface &o)
{
  unary(i, o, "Dilate iter", [&](pm_image *a, pm_image *b) { return pm_dilate_static(a, (int)s, iter, b, nullptr); });
}
void Erode(const IInterface &i, se::Compound_NL_SE s, const int size, IInterface &o)
{
  unary(i, o, "Erode compound", [&](pm_image *a, pm_image *b) { return pm_erode_compound(a, (int)s, size, b, nullptr); });
}
void Dilate(const IInterface &i, se::Compound_NL_SE s, const int size, IInterface &o)
{
  unary(i, o, "Dilate compound", [&](pm_image *a, pm_image *b) { return pm_dilate_compound(a, (int)s, size, b, nullptr); });
}

static std::vector<int64_t> nl_offsets(const se::IStructuringElement &s, std::size_t rank, int &n)
{
  auto nl = dynamic_cast<const se::neighbor_list *>(&s);
  if (!nl) fail("Unsupported structuring element type (neighbor_list only)");
  if (nl->rank() != rank) fail("SE and image have not the same rank");
  std::vector<int64_t> offs;
  for (const auto &c : nl->get_coordinates())
    for (auto v : c) offs.push_back((int64_t)v);
  n = (int)nl->GetSize();
  return offs;
}
void Erode(const IInterface &i, const se::IStructuringElement &s, IInterface &o)
{
  int n = 0;
  auto offs = nl_offsets(s, i.GetRank(), n);
  unary(i, o, "Erode", [&](pm_image *a, pm_image *b) { return pm_erode_nl(a, offs.data(), n, b, nullptr); });
}
void Dilate(const IInterface &i, const se::IStructuringElement &s, IInterface &o)
{
  int n = 0;
  auto offs = nl_offsets(s, i.GetRank(), n);
  unary(i, o, "Dilate", [&](pm_image *a, pm_image *b) { return pm_dilate_nl(a, offs.data(), n, b, nullptr); });
}

void ErodeX(const IInterface &i, const std::ptrdiff_t r, IInterface &o)
{
  unary(i, o, "ErodeX", [&](pm_image *a, pm_image *b) { return pm_erode_x(a, r, b, nullptr); });
}
void DilateX(const IInterface &i, const std::ptrdiff_t r, IInterface &o)
{
  unary(i, o, "DilateX", [&](pm_image *a, pm_image *b) { return pm_dilate_x(a, r, b, nullptr); });
}
// The reference's Y pass transposes into a temporary of the input's own shape and so
// rejects non-square images (ero_dil_line_se_t.hpp:258-259, transpose_op_t.hpp:54-60).
static void square_only(const IInterface &i, const char *what)
{
  auto s = i.GetShape();
  if (s.size() == 2 && s[0] != s[1]) fail(std::string(what) + ": ibd[0]!=obd[1] bound not compatible");
}
void ErodeY(const IInterface &i, const std::ptrdiff_t r, IInterface &o)
{
  square_only(i, "ErodeY");
  unary(i, o, "ErodeY", [&](pm_image *a, pm_image *b) { return pm_erode_y(a, r, b, nullptr); });
}
void DilateY(const IInterface &i, const std::ptrdiff_t r, IInterface &o)
{
  square_only(i, "DilateY");
  unary(i, o, "DilateY", [&](pm_image *a, pm_image *b) { return pm_dilate_y(a, r, b, nullptr); });
}
void Opening(const IInterface &i, se::Common_NL_SE s, const int iter, IInterface &o)
{
  unary(i, o, "Opening", [&](pm_image *a, pm_image *b) { return pm_opening(a, (int)s, iter, b, nullptr); });
}
void Closing(const IInterface &i, se::Common_NL_SE s, const int iter, IInterface &o)
{
  unary(i, o, "Closing", [&](pm_image *a, pm_image *b) { return pm_closing(a, (int)s, iter, b, nullptr); });
}
void Gradient(const IInterface &i, se::Common_NL_SE s, const int iter, IInterface &o)
{
  unary(i, o, "Gradient", [&](pm_image *a, pm_image *b) { return pm_gradient(a, (int)s, iter, b, nullptr); });
}

// ------------------------------------------------------------------ geodesy
namespace geo {

void Reconstruction(reconstruction_type t, const IInterface &marker, const IInterface &mask, se::Common_NL_SE s,
                    IInterface &o)
{
  binary(marker, mask, o, "Reconstruction", [&](pm_image *a, pm_image *b, pm_image *c) {
    return pm_reconstruct(t == reconstruction_type::dilate ? PM_REC_DILATE : PM_REC_ERODE, a, b, (int)s, c, nullptr);
  });
}
void low_leveling(const IInterface &r, const IInterface &m, se::Common_NL_SE s, IInterface &o)
{
  binary(r, m, o, "low_leveling", [&](pm_image *a, pm_image *b, pm_image *c) { return pm_low_leveling(a, b, (int)s, c, nullptr); });
}
void high_leveling(const IInterface &r, const IInterface &m, se::Common_NL_SE s, IInterface &o)
{
  binary(r, m, o, "high_leveling", [&](pm_image *a, pm_image *b, pm_image *c) { return pm_high_leveling(a, b, (int)s, c, nullptr); });
}
void leveling(const IInterface &r, const IInterface &m, se::Common_NL_SE s, IInterface &o)
{
  binary(r, m, o, "leveling", [&](pm_image *a, pm_image *b, pm_image *c) { return pm_leveling(a, b, (int)s, c, nullptr); });
}
#define PM_EXTREMA(NAME)                                                                                       \
  void NAME(const IInterface &i, const ScalarTypeVariant &v, se::Common_NL_SE s, IInterface &o)                \
  {                                                                                                            \
    const double h = scalar(v);                                                                                \
    unary(i, o, #NAME, [&](pm_image *a, pm_image *b) { return pm_##NAME(a, h, (int)s, b, nullptr); });        \
  }
PM_EXTREMA(h_minima)
PM_EXTREMA(h_maxima)
PM_EXTREMA(h_convex)
PM_EXTREMA(h_concave)
PM_EXTREMA(dynamic_pseudo_opening)
PM_EXTREMA(dynamic_pseudo_closing)
#undef PM_EXTREMA

}  // namespace geo

// ------------------------------------------------------------------ labelling
namespace label {

namespace {
// label_binary.cpp:21-35: the output must be a GINT64 image (downcast check), sizes
// compatible, images different
template <typename F> void labelled(const IInterface &i, IInterface &o, const std::string &what, F &&f)
{
  AssertSizesCompatible(i, o, what + " incompatible size");
  AssertImagesAreDifferent(i, o, what + " output must be != than input images");
  if (o.GetPType() != PType::PType_GrayINT64) fail(what + "ImageDispatch o_img downcast fail");
  Staged si(i, true), so(o, false);
  pm_check(f(si.desc(), so.desc()));
  so.download();
}
}  // namespace

std::size_t label_binary(const IInterface &i, se::Common_NL_SE s, IInterface &o)
{
  int64_t n = 0;
  labelled(i, o, "label_binary", [&](pm_image *a, pm_image *b) { return pm_label_binary(a, (int)s, b, &n, nullptr); });
  return (std::size_t)n;
}
std::size_t label_flat_zones(const IInterface &i, se::Common_NL_SE s, IInterface &o)
{
  int64_t n = 0;
  labelled(i, o, "label_flat_zones",
           [&](pm_image *a, pm_image *b) { return pm_label_flat_zones(a, (int)s, b, &n, nullptr); });
  return (std::size_t)n;
}
void label_minima(const IInterface &i, se::Common_NL_SE s, IInterface &o)
{
  labelled(i, o, "label_minima", [&](pm_image *a, pm_image *b) { return pm_label_minima(a, (int)s, b, nullptr); });
}
void label_maxima(const IInterface &i, se::Common_NL_SE s, IInterface &o)
{
  labelled(i, o, "label_maxima", [&](pm_image *a, pm_image *b) { return pm_label_maxima(a, (int)s, b, nullptr); });
}

}  // namespace label
}  // namespace poutre
