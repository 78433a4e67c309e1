// host_kats.cpp -- host-only checks of the C++ API (include/poutre_amd.hpp), no GPU needed:
//  * the reference's fixture-format KATs (test/base/image_iostring.cpp:28-242), read from a
//    cases file: ImageFromString must give the listed pixel type, compound type, shape and
//    pixel values; ImageToString of a filled image must give the listed string exactly;
//  * the reference's SE-model KATs (test/structuring_element/neighbor_list_se_t.cpp:20-69) on
//    the predefined se::details::neighbor_list_t objects;
//  * round trips and errors the KATs leave open (int64 extremes, rank 4, compound images given
//    to an operator, malformed strings).
//
// usage: host_kats <cases.txt>, one case per line:
//   read|<text>|<ptype>|<compound>|<rank>|<shape,...>|<values,...>
//   write|<ptype>|<compound>|<shape,...>|<fill,...>|<expected>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <functional>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "poutre_amd.hpp"

using namespace poutre;

static std::vector<std::string> split(const std::string &s, char c)
{
  std::vector<std::string> out;
  std::string cur;
  for (char ch : s) {
    if (ch == c) {
      out.push_back(cur);
      cur.clear();
    } else cur += ch;
  }
  out.push_back(cur);
  return out;
}

static std::vector<double> nums(const std::string &s)
{
  std::vector<double> v;
  for (const auto &t : split(s, ','))
    if (!t.empty()) v.push_back(std::stod(t));
  return v;
}

static PType ptype_of(const std::string &s)
{
  if (s == "GUINT8") return PType::PType_GrayUINT8;
  if (s == "GINT32") return PType::PType_GrayINT32;
  if (s == "GINT64") return PType::PType_GrayINT64;
  if (s == "F32") return PType::PType_F32;
  return PType::PType_D64;
}
static CompoundType ctype_of(const std::string &s)
{
  if (s == "3Planes") return CompoundType::CompoundType_3Planes;
  if (s == "4Planes") return CompoundType::CompoundType_4Planes;
  return CompoundType::CompoundType_Scalar;
}
static std::size_t planes(CompoundType c)
{
  return c == CompoundType::CompoundType_3Planes ? 3 : c == CompoundType::CompoundType_4Planes ? 4 : 1;
}

// value k of an image (k over pixels x planes), as a double
static double value_at(const IInterface &img, std::size_t k)
{
  const void *d = img.Data();
  switch (img.GetPType()) {
  case PType::PType_GrayUINT8: return static_cast<const pUINT8 *>(d)[k];
  case PType::PType_GrayINT32: return static_cast<const pINT32 *>(d)[k];
  case PType::PType_GrayINT64: return (double)static_cast<const pINT64 *>(d)[k];
  case PType::PType_F32: return static_cast<const pFLOAT *>(d)[k];
  default: return static_cast<const pDOUBLE *>(d)[k];
  }
}
static void set_at(IInterface &img, std::size_t k, double v)
{
  void *d = img.Data();
  switch (img.GetPType()) {
  case PType::PType_GrayUINT8: static_cast<pUINT8 *>(d)[k] = (pUINT8)v; break;
  case PType::PType_GrayINT32: static_cast<pINT32 *>(d)[k] = (pINT32)v; break;
  case PType::PType_GrayINT64: static_cast<pINT64 *>(d)[k] = (pINT64)v; break;
  case PType::PType_F32: static_cast<pFLOAT *>(d)[k] = (pFLOAT)v; break;
  default: static_cast<pDOUBLE *>(d)[k] = v; break;
  }
}

static int fails = 0, checks = 0;
static void expect(bool ok, const std::string &what)
{
  checks++;
  if (!ok) {
    fails++;
    std::printf("FAIL %s\n", what.c_str());
  }
}
static void expect_throw(const std::string &what, const std::function<void()> &fn)
{
  checks++;
  try {
    fn();
  } catch (const std::runtime_error &e) {
    if (std::strncmp(e.what(), "Throw ", 6) == 0) return;
    fails++;
    std::printf("FAIL %s: unexpected message %s\n", what.c_str(), e.what());
    return;
  }
  fails++;
  std::printf("FAIL %s: no exception\n", what.c_str());
}

static void run_read(const std::vector<std::string> &f)
{
  auto img = ImageFromString(f[1]);
  const std::string id = "read " + f[1].substr(0, 24);
  expect(img->GetPType() == ptype_of(f[2]), id + ": ptype");
  expect(img->GetCType() == ctype_of(f[3]), id + ": compound type");
  expect(img->GetRank() == (std::size_t)std::stoul(f[4]), id + ": rank");
  std::vector<std::size_t> shape;
  for (double d : nums(f[5])) shape.push_back((std::size_t)d);
  expect(img->GetShape() == shape, id + ": shape");
  const auto vals = nums(f[6]);
  std::size_t n = planes(img->GetCType());
  for (auto d : shape) n *= d;
  expect(vals.size() == n, id + ": value count");
  for (std::size_t k = 0; k < vals.size() && k < n; k++)
    if (value_at(*img, k) != vals[k]) {
      expect(false, id + ": value " + std::to_string(k));
      break;
    }
}

static void run_write(const std::vector<std::string> &f)
{
  std::vector<std::size_t> shape;
  for (double d : nums(f[3])) shape.push_back((std::size_t)d);
  const CompoundType c = ctype_of(f[2]);
  auto img = Create(shape, c, ptype_of(f[1]));
  const auto fill = nums(f[4]);
  std::size_t n = 1;
  for (auto d : shape) n *= d;
  for (std::size_t i = 0; i < n; i++)
    for (std::size_t p = 0; p < planes(c); p++) set_at(*img, i * planes(c) + p, fill[p % fill.size()]);
  const std::string got = ImageToString(*img);
  expect(got == f[5], "write " + f[5].substr(0, 24) + ": got [" + got + "]");
}

static void se_kats()
{
  // neighbor_list_se_t.cpp:20-26 remove center
  const auto sq = se::SESquare2D;
  const auto sq_nc = sq.remove_center();
  expect(sq_nc.is_symmetric(), "remove_center symmetric");
  expect(sq_nc.get_coordinates().size() == 8, "remove_center size 8");
  // :28-33 transpose
  expect(sq.is_equal_unordered(sq.transpose()), "transpose");
  // :35-39 symmetric
  expect(sq.is_symmetric(), "symmetric");
  // :41-52 split
  const auto [up, lo] = sq.split_upper_lower();
  expect(!up.is_symmetric() && !lo.is_symmetric(), "split 2D");
  const auto [up3, lo3] = se::SESquare3D.split_upper_lower();
  expect(!up3.is_symmetric() && !lo3.is_symmetric(), "split 3D");
  // :54-61 extension
  expect(sq.maximum_extension(0, true) == 1 && sq.maximum_extension(0, false) == 1 &&
             sq.maximum_extension(1, true) == 1 && sq.maximum_extension(1, false) == 1,
         "extension");
  // :63-69 extension pair
  const auto [ext_upper, ext_lower] = sq.maximum_extension();
  expect(ext_lower == poutre::details::av::idx2d{1, 1}, "extension pair lower");
  expect(ext_upper == poutre::details::av::idx2d{-1, -1}, "extension pair upper");

  // beyond the KATs: the split follows index<2>::operator< (last coordinate first), so the
  // upper half of the square is dx = -1 (three offsets) plus (dy, dx) = (-1, 0)
  using idx = poutre::details::av::idx2d;
  const std::vector<idx> want_up{{-1, -1}, {+0, -1}, {+1, -1}, {-1, +0}};
  expect(up.get_coordinates() == want_up, "split upper offsets");
  expect(lo.size() == 4 && up.size() == 4, "split sizes");
  expect(se::SESegmentX2D.maximum_extension(1, true) == 1 && se::SESegmentX2D.maximum_extension(0, true) == 0,
         "segment extension");
  expect(se::SECross3D.GetSize() == 7 && se::SESquare3D.GetSize() == 27, "3D sizes");
  // unordered equality compares the SETS of offsets (repeats do not count)
  const se::details::neighbor_list_t<2> rep{{0, 0}, {0, 0}, {-1, 0}}, once{{-1, 0}, {0, 0}};
  expect(rep.is_equal_unordered(once) && !rep.is_equal(once), "unordered equality");
  const se::details::neighbor_list_t<2> asym{{0, 0}, {1, 2}};
  expect(!asym.is_symmetric(), "asymmetric");
  expect_throw("split of an asymmetric SE", [&] { (void)asym.split_upper_lower(); });
  const se::details::neighbor_list_t<2> shifted{{1, 1}, {0, 1}};  // centre (first offset) (1, 1)
  const std::vector<idx> want_t{{1, 1}, {2, 1}};
  expect(shifted.transpose().get_coordinates() == want_t, "transpose about the first offset");
  // a runtime SE of another rank than the image is refused, before any device work
  auto img = Create({4, 4}, CompoundType::CompoundType_Scalar, PType::PType_GrayUINT8);
  auto out = Create({4, 4}, CompoundType::CompoundType_Scalar, PType::PType_GrayUINT8);
  expect_throw("SE rank", [&] { Dilate(*img, se::SESquare3D, *out); });
}

static void round_trips()
{
  const std::string i64 = "Scalar GINT64 1 3 9223372036854775807 -9223372036854775808 -1";
  expect(ImageToString(*ImageFromString(i64)) == i64, "int64 extremes round trip");
  const std::string u8 = "Scalar GUINT8 4 1 2 1 2 0 255 7 8";
  expect(ImageToString(*ImageFromString(u8)) == u8, "rank 4 round trip");
  const std::string f32 = "Scalar F32 1 3 0.1 -2.5 1e+10";
  expect(ImageToString(*ImageFromString(f32)) == f32, "float round trip (default stream format)");
  const std::string d64 = "4Planes D64 1 1 1.25 -3 0 7";
  expect(ImageToString(*ImageFromString(d64)) == d64, "4Planes round trip");
  expect_throw("unknown compound type", [] { (void)ImageFromString("Vector GUINT8 1 1 0"); });
  expect_throw("unknown pixel type", [] { (void)ImageFromString("Scalar GUINT9 1 1 0"); });
  expect_throw("NumDims 0", [] { (void)ImageFromString("Scalar GUINT8 0"); });
  expect_throw("missing values", [] { (void)ImageFromString("Scalar GINT32 1 3 1 2"); });
  expect_throw("not a number", [] { (void)ImageFromString("Scalar D64 1 2 1.5 x"); });
  // the operators take Scalar images only, and ranks 1..3
  auto c3 = ImageFromString("3Planes GUINT8 2 2 2 1 2 3 4 5 6 7 8 9 10 11 12");
  auto c3o = CloneGeometry(*c3);
  expect_throw("compound image to an operator", [&] { Dilate(*c3, se::Common_NL_SE::SESquare2D, 1, *c3o); });
  auto r4 = ImageFromString(u8);
  auto r4o = CloneGeometry(*r4);
  expect_throw("rank 4 image to an operator", [&] { Dilate(*r4, se::Common_NL_SE::SESquare2D, 1, *r4o); });
}

int main(int argc, char **argv)
{
  if (argc < 2) return 2;
  std::ifstream in(argv[1]);
  std::string line;
  int cases = 0;
  while (std::getline(in, line)) {
    if (line.empty()) continue;
    const auto f = split(line, '|');
    try {
      if (f[0] == "read" && f.size() == 7) run_read(f);
      else if (f[0] == "write" && f.size() == 6) run_write(f);
      else return 3;
      cases++;
    } catch (const std::exception &e) {
      fails++;
      std::printf("FAIL %s: exception %s\n", f[0].c_str(), e.what());
    }
  }
  se_kats();
  round_trips();
  std::printf("host_kats: %d fixture cases, %d checks, %d failures\n", cases, checks, fails);
  return fails ? 1 : 0;
}
