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

static std::size_t planes_of(CompoundType c)
{
  switch (c) {
  case CompoundType::CompoundType_Scalar: return 1;
  case CompoundType::CompoundType_3Planes: return 3;
  case CompoundType::CompoundType_4Planes: return 4;
  default: fail("Unsupported compound type");
  }
}

HostImage::HostImage(std::vector<std::size_t> shape, PType ptype, CompoundType ctype)
    : m_shape(std::move(shape)), m_ptype(ptype), m_ctype(ctype),
      m_data(count(m_shape) * planes_of(ctype) * SizeOf(ptype))
{
  if (m_shape.empty() || m_shape.size() > 4) fail("Unsupported number of dims");
}
std::string HostImage::str() const { return describe("HostImage", m_shape, m_ptype); }
std::size_t HostImage::size() const { return count(m_shape); }
std::size_t HostImage::planes() const { return planes_of(m_ctype); }

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
  const std::size_t bytes = count(src.GetShape()) * planes_of(src.GetCType()) * SizeOf(src.GetPType());
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
  (void)planes_of(ctype);  // Scalar, 3Planes or 4Planes
  (void)SizeOf(ptype);
  return std::make_unique<HostImage>(dims, ptype, ctype);
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
// ImageFromString / ImageToString (src/base/image_interface.cpp:616-909, type names
// src/base/types.cpp:201-345): values go through each PType's text type (TypeTraits::str_type,
// types_traits.hpp:118-243) and the default stream format.
static PType ptype_from(const std::string &s)
{
  if (s == "GUINT8") return PType::PType_GrayUINT8;
  if (s == "GUINT16") return PType::PType_GrayUINT16;
  if (s == "GINT32") return PType::PType_GrayINT32;
  if (s == "GINT64") return PType::PType_GrayINT64;
  if (s == "F32") return PType::PType_F32;
  if (s == "D64") return PType::PType_D64;
  fail("Unable to read PType from stream");
}
static CompoundType ctype_from(const std::string &s)
{
  if (s == "Scalar") return CompoundType::CompoundType_Scalar;
  if (s == "3Planes") return CompoundType::CompoundType_3Planes;
  if (s == "4Planes") return CompoundType::CompoundType_4Planes;
  fail("Unable to read CompoundType from stream");
}
static const char *ctype_name(CompoundType c)
{
  switch (c) {
  case CompoundType::CompoundType_Scalar: return "Scalar";
  case CompoundType::CompoundType_3Planes: return "3Planes";
  case CompoundType::CompoundType_4Planes: return "4Planes";
  default: return "Unknown PType";
  }
}

// T: storage type, S: the text type values are read as and written as
template <typename T, typename S> static void read_values(std::istringstream &is, void *data, std::size_t n,
                                                          const std::string &src)
{
  T *p = static_cast<T *>(data);
  for (std::size_t i = 0; i < n; i++) {
    S v{};
    is >> v;
    if (is.fail()) fail("ImageFromStringDenseScalarDispatchPType: unable to read current pixel from " + src);
    p[i] = static_cast<T>(v);
  }
}
template <typename T, typename S> static void write_values(std::ostringstream &os, const void *data, std::size_t n)
{
  const T *p = static_cast<const T *>(data);
  for (std::size_t i = 0; i < n; i++) os << static_cast<S>(p[i]) << ' ';
}

std::unique_ptr<IInterface> ImageFromString(const std::string &str)
{
  std::istringstream is(str);
  if (!is.good()) fail("ImageFromString: unable to read input string");
  std::string cs, ps;
  is >> cs;
  if (is.fail()) fail("ImageFromString: unable to read CompoundType from " + str);
  const CompoundType c = ctype_from(cs);
  is >> ps;
  if (is.fail()) fail("ImageFromString: unable to read PType from " + str);
  const PType p = ptype_from(ps);
  std::ptrdiff_t rank = -1;
  is >> rank;
  if (is.fail()) fail("ImageFromString: unable to read NumDims from " + str);
  if (!(rank > 0)) fail("ImageFromString: wrong value from NumDims " + str);
  std::vector<std::size_t> shape;
  for (std::ptrdiff_t i = 0; i < rank; i++) {
    std::ptrdiff_t d = 0;
    is >> d;
    if (is.fail()) fail("ImageFromString: unable to read dim from " + str);
    shape.push_back(static_cast<std::size_t>(d));
  }
  if (rank > 4) fail("Unsupported number of dims");
  auto img = Create(shape, c, p);
  const std::size_t n = count(shape) * planes_of(c);
  switch (p) {
  case PType::PType_GrayUINT8: read_values<pUINT8, unsigned int>(is, img->Data(), n, str); break;
  case PType::PType_GrayUINT16: read_values<pUINT16, unsigned int>(is, img->Data(), n, str); break;
  case PType::PType_GrayINT32: read_values<pINT32, pINT32>(is, img->Data(), n, str); break;
  case PType::PType_GrayINT64: read_values<pINT64, pINT64>(is, img->Data(), n, str); break;
  case PType::PType_F32: read_values<pFLOAT, pFLOAT>(is, img->Data(), n, str); break;
  case PType::PType_D64: read_values<pDOUBLE, pDOUBLE>(is, img->Data(), n, str); break;
  default: fail("ImageFromStringDenseScalarDispatchPType::Unsupported ptype");
  }
  return img;  // trailing values, if any, are ignored as by the reference's parser
}

std::string ImageToString(const IInterface &i_img)
{
  std::unique_ptr<IInterface> host;
  const IInterface *src = &i_img;
  if (i_img.OnDevice()) {
    host = Create(i_img.GetShape(), i_img.GetCType(), i_img.GetPType());
    CopyInto(i_img, *host);
    src = host.get();
  }
  if (src->GetRank() == 0) fail("Unsupported number of dims:0");
  std::ostringstream os;
  os << ctype_name(src->GetCType()) << ' ' << ptype_name(src->GetPType()) << ' ' << src->GetRank() << ' ';
  for (auto d : src->GetShape()) os << d << ' ';
  const std::size_t n = count(src->GetShape()) * planes_of(src->GetCType());
  switch (src->GetPType()) {
  case PType::PType_GrayUINT8: write_values<pUINT8, unsigned int>(os, src->Data(), n); break;
  case PType::PType_GrayUINT16: write_values<pUINT16, unsigned int>(os, src->Data(), n); break;
  case PType::PType_GrayINT32: write_values<pINT32, pINT32>(os, src->Data(), n); break;
  case PType::PType_GrayINT64: write_values<pINT64, pINT64>(os, src->Data(), n); break;
  case PType::PType_F32: write_values<pFLOAT, pFLOAT>(os, src->Data(), n); break;
  case PType::PType_D64: write_values<pDOUBLE, pDOUBLE>(os, src->Data(), n); break;
  default: fail("ImageToStringDenseScalarDispatchPType::Unsupported pixel type");
  }
  std::string out = os.str();
  out.pop_back();  // the separator after the last field
  return out;
}

// ------------------------------------------------------------------ structuring elements
namespace se {

// predefined runtime SEs: offsets and their order as the reference lists them
// (src/structuring_element/predefined_nl_se.cpp:19-91), {dy, dx} / {dz, dy, dx}
const details::neighbor_list_t<1> SESegmentX1D{{0}, {-1}, {+1}};
const details::neighbor_list_t<2> SESegmentX2D{{0, 0}, {0, -1}, {0, +1}};
const details::neighbor_list_t<2> SESegmentY2D{{0, 0}, {-1, 0}, {+1, 0}};
const details::neighbor_list_t<2> SESquare2D{{0, 0},   {-1, -1}, {+0, -1}, {+1, -1}, {-1, +0},
                                             {+1, +0}, {-1, +1}, {+0, +1}, {+1, +1}};
const details::neighbor_list_t<2> SECross2D{{0, 0}, {+0, -1}, {-1, +0}, {+1, +0}, {+0, +1}};
const details::neighbor_list_t<3> SECross3D{{0, 0, 0},   {+0, +0, -1}, {+0, -1, +0}, {-1, +0, +0},
                                            {+1, +0, +0}, {+0, +1, +0}, {+0, +0, +1}};
const details::neighbor_list_t<3> SESquare3D{
    {0, 0, 0},    {-1, -1, -1}, {+0, -1, -1}, {+1, -1, -1}, {-1, +0, -1}, {+0, +0, -1}, {+1, +0, -1},
    {-1, +1, -1}, {+0, +1, -1}, {+1, +1, -1}, {-1, -1, +0}, {+0, -1, +0}, {+1, -1, +0}, {-1, +0, +0},
    {+1, +0, +0}, {-1, +1, +0}, {+0, +1, +0}, {+1, +1, +0}, {-1, -1, +1}, {+0, -1, +1}, {+1, -1, +1},
    {-1, +0, +1}, {+0, +0, +1}, {+1, +0, +1}, {-1, +1, +1}, {+0, +1, +1}, {+1, +1, +1}};
const details::neighbor_list_t<3> SESegmentX3D{{0, 0, 0}, {0, 0, -1}, {0, 0, +1}};
const details::neighbor_list_t<3> SESegmentY3D{{0, 0, 0}, {0, -1, 0}, {0, +1, 0}};
const details::neighbor_list_t<3> SESegmentZ3D{{0, 0, 0}, {-1, 0, 0}, {+1, 0, 0}};

}  // namespace se

// ------------------------------------------------------------------ staging
namespace {

// Device view of an image: the image itself when it lives in HBM, a staged copy otherwise.
class Staged
{
public:
  Staged(const IInterface &img, bool upload) : m_img(img)
  {
    if (img.GetCType() != CompoundType::CompoundType_Scalar) fail("Unsupported compound type");
    if (img.GetRank() == 0 || img.GetRank() > 3) fail("Unsupported number of dims");
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

// the C ABI carries scalars as double: an int64 alternative beyond 2^53 would be rounded, so
// it is refused instead
double scalar(const ScalarTypeVariant &v)
{
  if (const auto *i = std::get_if<pINT64>(&v))
    if (*i > (pINT64(1) << 53) || *i < -(pINT64(1) << 53))
      fail("scalar operand " + std::to_string(*i) + " is not exactly representable (|value| > 2^53)");
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

// the offsets of a runtime SE whose rank is the image's (ero_dil.cpp:59-243: a
// neighbor_list_t of the image's rank, anything else is refused)
template <std::ptrdiff_t R> static bool nl_of_rank(const se::IStructuringElement &s, std::vector<int64_t> &offs, int &n)
{
  const auto *nl = dynamic_cast<const se::details::neighbor_list_t<R> *>(&s);
  if (!nl) return false;
  for (const auto &c : nl->get_coordinates())
    for (std::size_t k = 0; k < (std::size_t)R; k++) offs.push_back((int64_t)c[k]);
  n = (int)nl->size();
  return true;
}
static std::vector<int64_t> nl_offsets(const se::IStructuringElement &s, std::size_t rank, int &n, const char *what)
{
  std::vector<int64_t> offs;
  bool ok = false;
  switch (rank) {
  case 0: fail(std::string(what) + " Unsupported number of dims:0");
  case 1: ok = nl_of_rank<1>(s, offs, n); break;
  case 2: ok = nl_of_rank<2>(s, offs, n); break;
  case 3: ok = nl_of_rank<3>(s, offs, n); break;
  default: fail(std::string(what) + " Unsupported number of dims");
  }
  if (!ok) fail(std::string(what) + " Unsupported IStructuringElement type");
  return offs;
}
void Erode(const IInterface &i, const se::IStructuringElement &s, IInterface &o)
{
  checks(i, o, "Erode");  // the reference's checks come before the rank / SE dispatch
  int n = 0;
  auto offs = nl_offsets(s, i.GetRank(), n, "Erode");
  unary(i, o, "Erode", [&](pm_image *a, pm_image *b) { return pm_erode_nl(a, offs.data(), n, b, nullptr); });
}
void Dilate(const IInterface &i, const se::IStructuringElement &s, IInterface &o)
{
  checks(i, o, "Dilate");  // the reference's checks come before the rank / SE dispatch
  int n = 0;
  auto offs = nl_offsets(s, i.GetRank(), n, "Dilate");
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
