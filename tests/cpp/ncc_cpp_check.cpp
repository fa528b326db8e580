// Checks of the C++ layer (include/mv_ncc.hpp) over the product library,
// written the way a user of mvkit's C++ API calls it (examples/ncc_matching.cpp
// procedural and object styles; include/mvkit/cpp/mv_ncc_matching.hpp).
//
//   ncc_cpp_check host <template.u8> <w> <h> <workdir>
//     tuples, handles, models, persistence, errors: no device work.  Prints
//     "params <levels> <start> <extent> <step> <metric>" and "ok <check>"
//     lines; exits 1 at the first failed check.
//   ncc_cpp_check find <template.u8> <w> <h> <image.u8> <w> <h> <start> <extent> <min_score> <num_matches>
//     FindNccModel through the object, procedural and image-member forms
//     (which must agree exactly) and FindNccModels over two models; prints
//     "find r c a s" and "models r c a s m" lines (%.17g).
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iterator>
#include <string>
#include <vector>

#include "mv_ncc.hpp"

namespace {

void expect(bool ok, const char* what) {
  if (!ok) {
    std::fprintf(stderr, "FAILED: %s\n", what);
    std::exit(1);
  }
}
void pass(const char* what) { std::printf("ok %s\n", what); }

std::string str(::MvStringView v) { return std::string(v.data ? v.data : "", v.size); }

// Runs `f`, which must throw an MvException with this status, code and (if
// given) operator and parameter names.
template <class F>
void expect_throw(F f, mv::MvStatus status, mv::MvErrorCode code, const char* op, const char* param,
                  const char* what) {
  try {
    f();
  } catch (const mv::MvException& e) {
    const bool ok = e.GetStatus() == status && e.GetErrorCode() == code &&
                    (!op || str(e.GetOperatorName()) == op) && (!param || str(e.GetParameterName()) == param);
    if (!ok)
      std::fprintf(stderr, "%s: got status %d code %d op '%s' param '%s' (%s)\n", what, int(e.GetStatus()),
                   int(e.GetErrorCode()), str(e.GetOperatorName()).c_str(), str(e.GetParameterName()).c_str(),
                   e.what());
    expect(ok, what);
    return;
  }
  expect(false, what);
}

std::vector<unsigned char> read_file(const char* path) {
  std::ifstream f(path, std::ios::binary);
  expect(bool(f), "input file opens");
  return std::vector<unsigned char>(std::istreambuf_iterator<char>(f), {});
}

mv::MvImage load(const char* path, long w, long h, std::vector<unsigned char>& pixels) {
  pixels = read_file(path);
  expect(pixels.size() == size_t(w) * size_t(h), "input file has w*h bytes");
  mv::MvImage im;
  im.GenImage1("byte", w, h, pixels.data());
  return im;
}

void tuples() {
  mv::MvTuple t(1);
  t.Append(2.5).Append("abc");
  expect(t.Length() == 3 && t.I(0) == 1 && t.D(1) == 2.5 && t.S(2) == mv::MvString("abc"), "tuple values");
  expect(t.D(0) == 1.0, "integer read as real");
  expect(t.Kind(0) == MV_VALUE_INT64 && t.Kind(1) == MV_VALUE_REAL64 && t.Kind(2) == MV_VALUE_STRING, "kinds");
  expect_throw([&] { (void)t.I(1); }, mv::MvStatus::TypeMismatch, mv::MvErrorCode::InvalidType, "get_i", nullptr,
               "real read as integer");
  expect_throw([&] { (void)t.I(9); }, mv::MvStatus::OutOfRange, mv::MvErrorCode::IndexOutOfRange, nullptr, nullptr,
               "index out of range");
  mv::MvTuple c(t);
  c.Append(int64_t(4));
  expect(t.Length() == 3 && c.Length() == 4, "copies are independent");
  mv::MvTuple moved(std::move(c));
  expect(moved.Length() == 4 && c.Length() == 0, "move leaves the empty tuple");
  const mv::MvTuple sel = moved.TupleSelect(mv::MvTuple(std::vector<int64_t>{3, 0}));
  expect(sel.Length() == 2 && sel.I(0) == 4 && sel.I(1) == 1, "TupleSelect");
  const mv::MvTuple range = moved.TupleSelectRange(1, 2);
  expect(range.Length() == 2 && range.D(0) == 2.5 && range.S(1) == mv::MvString("abc"), "TupleSelectRange");
  const mv::MvTuple cat = range.TupleConcat(sel);
  expect(cat.Length() == 4 && cat.I(3) == 1, "TupleConcat");
  const mv::MvTuple empty;
  expect(empty.Length() == 0, "default tuple is empty");
  expect_throw([&] { (void)empty.I(0); }, mv::MvStatus::OutOfRange, mv::MvErrorCode::IndexOutOfRange, nullptr,
               nullptr, "read from the empty tuple");
  mv::MvTuple sem;
  mv::TupleSemType(mv::MvTuple(std::vector<double>{1.0, 2.0}), &sem);
  expect(sem.S() == mv::MvString("real"), "sem type real");
  mv::TupleSemType(t, &sem);
  expect(sem.S() == mv::MvString("any"), "sem type mixed");
  mv::TupleSemType(empty, &sem);
  expect(sem.S() == mv::MvString("any"), "sem type empty");
  pass("tuples");
}

void params_equal(const mv::MvNCCModel& a, const mv::MvNCCModel& b, const char* what) {
  double s1, e1, st1, s2, e2, st2;
  mv::MvString m1, m2;
  const int64_t l1 = a.GetNccModelParams(&s1, &e1, &st1, &m1), l2 = b.GetNccModelParams(&s2, &e2, &st2, &m2);
  expect(l1 == l2 && s1 == s2 && e1 == e2 && st1 == st2 && m1 == m2, what);
}

void host(const char* tpath, long tw, long th, const std::string& dir) {
  tuples();
  std::vector<unsigned char> tp;
  const mv::MvImage templ = load(tpath, tw, th, tp);
  const double start = -0.35, extent = 0.7;

  // object style (examples/ncc_matching.cpp:42-43)
  mv::MvNCCModel model;
  expect(!model.IsInitialized() && model.Type() == mv::MvObjectType::Unknown, "default model is uninitialized");
  model.CreateNccModel(templ, "auto", start, extent, "auto", "use_polarity");
  expect(model.IsInitialized() && model.Type() == mv::MvObjectType::NccModel, "created model is an NCC model");
  double s, e, st;
  mv::MvString metric;
  const int64_t levels = model.GetNccModelParams(&s, &e, &st, &metric);
  std::printf("params %lld %.17g %.17g %.17g %s\n", (long long)levels, s, e, st, metric.Text());
  expect(s == start && e == extent && metric == mv::MvString("use_polarity") && levels >= 1, "params");

  // procedural style (examples/ncc_matching.cpp:16): the same model
  mv::MvTuple id, L, S, E, St, M;
  mv::CreateNccModel(templ, "auto", start, extent, "auto", "use_polarity", &id);
  mv::GetNccModelParams(id, &L, &S, &E, &St, &M);
  expect(L.I() == levels && S.D() == s && E.D() == e && St.D() == st && M.S() == metric, "procedural params");
  const mv::MvNCCModel typed(id.H());
  params_equal(model, typed, "typed handle");
  const mv::MvNCCModel from_image = templ.CreateNccModel(int64_t(levels), start, extent, st, "use_polarity");
  params_equal(model, from_image, "MvImage::CreateNccModel");
  pass("create");

  // errors: controls (create_ncc_model, ncc_matching_api.cpp:29-60) and the
  // C++ layer's own output checks (api_bridge.hpp)
  expect_throw([&] { mv::MvNCCModel m; m.CreateNccModel(templ, int64_t(17), 0.0, 0.0, 0.0, "use_polarity"); },
               mv::MvStatus::InvalidArgument, mv::MvErrorCode::InvalidValue, "create_ncc_model", "NumLevels",
               "NumLevels 17");
  expect_throw([&] { mv::MvNCCModel m; m.CreateNccModel(templ, "auto", 0.0, 0.0, "auto", "sad"); },
               mv::MvStatus::InvalidArgument, mv::MvErrorCode::UnsupportedValue, "create_ncc_model", "Metric",
               "unknown metric");
  expect_throw([&] { mv::CreateNccModel(templ, "auto", 0.0, 0.0, "auto", "use_polarity", nullptr); },
               mv::MvStatus::InvalidArgument, mv::MvErrorCode::NullArgument, "CreateNccModel", "output",
               "null output");
  expect_throw([&] { mv::GetNccModelParams(id, &L, &L, &E, &St, &M); }, mv::MvStatus::InvalidArgument,
               mv::MvErrorCode::InvalidValue, "GetNccModelParams", "Output", "repeated output");
  expect_throw([&] { (void)model.GetNccModelParams(&s, &s, &st, &metric); }, mv::MvStatus::InvalidArgument,
               mv::MvErrorCode::InvalidValue, "GetNccModelParams", "Output", "overlapping member outputs");
  mv::MvTuple r, c, a, q, mi;
  expect_throw([&] { model.FindNccModel(templ, start, extent, 1.5, 1, 0.5, "true", 0, &r, &c, &a, &q); },
               mv::MvStatus::InvalidArgument, mv::MvErrorCode::InvalidValue, "find_ncc_model", "MinScore",
               "MinScore 1.5");
  expect_throw([&] { model.FindNccModel(templ, start, extent, 0.8, 1, 0.5, "maybe", 0, &r, &c, &a, &q); },
               mv::MvStatus::InvalidArgument, mv::MvErrorCode::UnsupportedValue, "find_ncc_model", "SubPixel",
               "SubPixel keyword");
  expect_throw([&] { mv::FindNccModel(templ, id, start, extent, 0.8, 1, 0.5, "true", 0, &r, &c, &r, &q); },
               mv::MvStatus::InvalidArgument, mv::MvErrorCode::InvalidValue, "FindNccModel", "Output",
               "repeated find output");
  // an empty image gives empty outputs (no device work)
  const mv::MvImage nothing;
  model.FindNccModel(nothing, start, extent, 0.8, 1, 0.5, "true", 0, &r, &c, &a, &q);
  expect(r.Length() == 0 && c.Length() == 0 && a.Length() == 0 && q.Length() == 0, "empty image");
  pass("errors");

  // persistence: file (persist.cpp:14-84) and serialized items (serialization_api.cpp:51-141)
  const std::string path = dir + "/model.ncm";
  model.WriteNccModel(path.c_str());
  mv::MvNCCModel read;
  read.ReadNccModel(path.c_str());
  params_equal(model, read, "read model");
  mv::MvTuple read_id;
  mv::ReadNccModel(path.c_str(), &read_id);
  expect_throw([&] { mv::ReadNccModel((dir + "/missing.ncm").c_str(), &read_id); }, mv::MvStatus::IoError,
               mv::MvErrorCode::IoFailure, "read_ncc_model", nullptr, "missing file");
  const mv::MvSerializedItem item = model.SerializeHandle();
  expect(item.Type() == mv::MvObjectType::SerializedItem, "serialized item type");
  int64_t bytes = 0;
  void* p = item.GetSerializedItemPtr(&bytes);
  expect(p && bytes > 0, "serialized bytes");
  const mv::MvSerializedItem copy(p, bytes, "true");
  mv::MvNCCModel restored;
  restored.DeserializeHandle(copy);
  params_equal(model, restored, "deserialized model");
  expect_throw([&] { mv::MvNCCModel m{mv::MvHandle(item)}; }, mv::MvStatus::TypeMismatch,
               mv::MvErrorCode::InvalidType, "MvNCCModel", "ModelID", "item is not a model");
  expect_throw([&] { mv::MvSerializedItem x{mv::MvHandle(model)}; }, mv::MvStatus::TypeMismatch,
               mv::MvErrorCode::InvalidType, "MvSerializedItem", nullptr, "model is not an item");
  pass("persistence");

  // model arrays (mv_ncc_matching.hpp:262-380)
  const mv::MvNCCModel pair[2] = {model, read};
  mv::MvNCCModelArray arr(pair, 2);
  expect(arr.Length() == 2 && arr.Tools() != nullptr, "array length");
  const mv::MvTuple ids = arr.ConvertToTuple();
  mv::MvTuple sem;
  mv::TupleSemType(ids, &sem);
  expect(ids.Length() == 2 && sem.S() == mv::MvString("ncc_model"), "array tuple");
  mv::MvNCCModelArray shared = arr;
  arr.Clear();
  expect(arr.Length() == 0 && shared.Length() == 2, "array copies share, Clear drops one reference");
  mv::MvNCCModelArray set;
  set.SetFromTuple(ids);
  expect(set.Length() == 2, "SetFromTuple");
  expect_throw([&] { set.SetFromTuple(mv::MvTuple(item)); }, mv::MvStatus::TypeMismatch,
               mv::MvErrorCode::InvalidType, "MvNCCModel", nullptr, "array of items");
  expect(set.Length() == 2, "failed SetFromTuple leaves the array");
  expect_throw([&] { mv::MvNCCModelArray bad(pair, -1); }, mv::MvStatus::InvalidArgument,
               mv::MvErrorCode::InvalidValue, "MvNCCModelArray", "models", "negative length");
  pass("arrays");

  // regions and domains
  mv::MvRegion rect;
  rect.GenRectangle1(2, 3, 5, 9);
  mv::MvTuple rows, cb, ce;
  rect.GetRegionRuns(&rows, &cb, &ce);
  expect(rows.Length() == 4 && rows.I(0) == 2 && cb.I(0) == 3 && ce.I(3) == 9, "rectangle runs");
  const mv::MvImage reduced = templ.ReduceDomain(rect);
  mv::MvTuple rows2, cb2, ce2;
  reduced.GetDomain().GetRegionRuns(&rows2, &cb2, &ce2);
  expect(rows2.Length() == 4 && rows2.I(3) == 5 && cb2.I(2) == 3 && ce2.I(1) == 9, "reduced domain");
  pass("domains");

  // handle sharing: clearing one reference clears every copy
  mv::MvNCCModel alias = model;
  const mv::MvTuple alias_id(alias);
  mv::ClearHandle(alias_id);
  expect(!model.IsInitialized() && !alias.IsInitialized(), "cleared everywhere");
  expect_throw([&] { (void)model.GetNccModelParams(&s, &e, &st, &metric); }, mv::MvStatus::InvalidState,
               mv::MvErrorCode::InvalidObjectState, nullptr, nullptr, "cleared model");
  expect_throw([&] { mv::MvNCCModel m{mv::MvHandle(model)}; }, mv::MvStatus::TypeMismatch,
               mv::MvErrorCode::InvalidType, "MvNCCModel", "ModelID", "cleared handle is not a model");
  params_equal(read, restored, "other models unaffected");
  pass("clear");
  std::printf("all ok\n");
}

void print_rows(const char* tag, const mv::MvTuple& r, const mv::MvTuple& c, const mv::MvTuple& a,
                const mv::MvTuple& s, const mv::MvTuple* m) {
  for (size_t i = 0; i < r.Length(); ++i) {
    std::printf("%s %.17g %.17g %.17g %.17g", tag, r.D(i), c.D(i), a.D(i), s.D(i));
    if (m) std::printf(" %lld", (long long)m->I(i));
    std::printf("\n");
  }
}

bool same(const mv::MvTuple& x, const mv::MvTuple& y) {
  if (x.Length() != y.Length()) return false;
  for (size_t i = 0; i < x.Length(); ++i)
    if (x.D(i) != y.D(i)) return false;
  return true;
}

void find(char** argv) {
  std::vector<unsigned char> tp, ip;
  const mv::MvImage templ = load(argv[2], std::atol(argv[3]), std::atol(argv[4]), tp);
  const mv::MvImage image = load(argv[5], std::atol(argv[6]), std::atol(argv[7]), ip);
  const double start = std::atof(argv[8]), extent = std::atof(argv[9]), min_score = std::atof(argv[10]);
  const int64_t n = std::atoll(argv[11]);

  mv::MvNCCModel model;
  model.CreateNccModel(templ, int64_t(0), start, extent, "auto", "use_polarity");
  mv::MvTuple r, c, a, s;
  model.FindNccModel(image, start, extent, min_score, n, 0.5, "true", 0, &r, &c, &a, &s);
  print_rows("find", r, c, a, s, nullptr);

  mv::MvTuple r2, c2, a2, s2;
  mv::FindNccModel(image, mv::MvTuple(model), start, extent, min_score, n, 0.5, "true", 0, &r2, &c2, &a2, &s2);
  expect(same(r, r2) && same(c, c2) && same(a, a2) && same(s, s2), "procedural find equals object find");
  image.FindNccModel(model, start, extent, min_score, n, 0.5, "true", 0, &r2, &c2, &a2, &s2);
  expect(same(r, r2) && same(c, c2) && same(a, a2) && same(s, s2), "image member find equals object find");

  // two models (the second restored from the first's serialized bytes)
  mv::MvNCCModel second;
  second.DeserializeHandle(model.SerializeHandle());
  const mv::MvNCCModel both[2] = {model, second};
  mv::MvTuple r3, c3, a3, s3, m3;
  mv::MvNCCModel::FindNccModels(image, mv::MvNCCModelArray(both, 2), start, extent, min_score, n, 0.5, "true", 0,
                                &r3, &c3, &a3, &s3, &m3);
  print_rows("models", r3, c3, a3, s3, &m3);
  mv::MvTuple r4, c4, a4, s4, m4;
  model.FindNccModels(image, start, extent, min_score, n, 0.5, "true", 0, &r4, &c4, &a4, &s4, &m4);
  expect(same(r, r4) && same(s, s4) && m4.Length() == r.Length(), "one model through FindNccModels");
  std::printf("all ok\n");
}

}  // namespace

int main(int argc, char** argv) {
  try {
    if (argc == 6 && std::string(argv[1]) == "host") {
      host(argv[2], std::atol(argv[3]), std::atol(argv[4]), argv[5]);
      return 0;
    }
    if (argc == 12 && std::string(argv[1]) == "find") {
      find(argv);
      return 0;
    }
  } catch (const mv::MvException& e) {
    std::fprintf(stderr, "unexpected MvException: %s\n", e.what());
    return 1;
  }
  std::fprintf(stderr, "usage: ncc_cpp_check host <t.u8> <w> <h> <dir> | find <t.u8> <w> <h> <i.u8> <w> <h> "
                       "<start> <extent> <min_score> <num_matches>\n");
  return 2;
}
