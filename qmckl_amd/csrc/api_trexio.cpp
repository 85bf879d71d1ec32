// TREXIO input: qmckl_trexio_read (org/qmckl_trexio.org:1104-1190) fills the electron, nucleus,
// AO-basis and MO-basis groups of a context from a TREXIO file through the public setters, in the
// reference's order and with its reformatting of the per-shell and per-primitive index arrays
// (org/qmckl_trexio.org:400-800).
//
// libtrexio is not linked.  The TEXT back end, a directory with one "<group>.txt" file per group,
// is parsed here; a path that is not such a directory (an HDF5 file, a missing path) fails with
// QMCKL_INVALID_ARG_2 as the reference does when no back end opens it (org/qmckl_trexio.org:88-116).
//
// A group file starts with a header of "<key> <value>..." lines: scalars ("nucleus_num 2"), their
// "<name>_isSet" flags, "rank_<array> r", "dims_<array> <axis> <extent>" and "len_<string> n".
// After the header every array or string follows as a line holding its name alone and then its
// values, one per line, in C order of the stored dimensions.
#include <sys/stat.h>

#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "internal.hpp"

namespace {

struct Group {
  std::map<std::string, std::vector<std::string>> head;  // header key -> its remaining tokens
  std::map<std::string, std::map<int, int64_t>> dims;    // array -> axis -> extent
  std::vector<std::string> body;                         // lines after the header
  std::string name;
};

std::vector<std::string> tokens(const std::string& line) {
  std::istringstream is(line);
  std::vector<std::string> t;
  std::string w;
  while (is >> w) t.push_back(w);
  return t;
}

bool load_group(const std::string& dir, const char* group, Group& g, std::string& err) {
  g.name = group;
  const std::string path = dir + "/" + group + ".txt";
  std::ifstream f(path);
  if (!f) {
    err = "cannot open " + path;
    return false;
  }
  std::string line;
  bool in_head = true;
  while (std::getline(f, line)) {
    if (!line.empty() && line.back() == '\r') line.pop_back();
    if (in_head) {
      std::vector<std::string> t = tokens(line);
      if (t.size() >= 2) {
        if (t[0].compare(0, 5, "dims_") == 0) {
          if (t.size() != 3) {
            err = path + ": malformed line '" + line + "'";
            return false;
          }
          g.dims[t[0].substr(5)][std::atoi(t[1].c_str())] = std::strtoll(t[2].c_str(), nullptr, 10);
        } else {
          g.head[t[0]] = std::vector<std::string>(t.begin() + 1, t.end());
        }
        continue;
      }
      in_head = false;
    }
    g.body.push_back(line);
  }
  return true;
}

bool parse_int(const std::string& s, int64_t& v) {
  errno = 0;
  char* end = nullptr;
  v = std::strtoll(s.c_str(), &end, 10);
  return errno == 0 && end != s.c_str() && *end == '\0';
}

bool parse_double(const std::string& s, double& v) {
  errno = 0;
  char* end = nullptr;
  v = std::strtod(s.c_str(), &end);
  return errno == 0 && end != s.c_str() && *end == '\0';
}

// A scalar item: its "<name>_isSet" flag is 1 (when present) and its header line holds one integer.
bool read_scalar(const Group& g, const char* name, int64_t& v, std::string& err) {
  const std::string key(name);
  auto set = g.head.find(key + "_isSet");
  auto it = g.head.find(key);
  if ((set != g.head.end() && (set->second.empty() || set->second[0] != "1")) || it == g.head.end() ||
      it->second.size() != 1 || !parse_int(it->second[0], v)) {
    err = g.name + ".txt: " + key + " is not set";
    return false;
  }
  return true;
}

// The lines of an array item: its rank and extents from the header, `expected` values after the
// line holding its name.  Returns false with err empty when the file does not hold the array.
bool array_lines(const Group& g, const char* name, int64_t expected, std::vector<std::string>& out,
                 std::string& err) {
  const std::string key(name);
  auto rk = g.head.find("rank_" + key);
  if (rk == g.head.end()) return false;
  int64_t rank = 0;
  if (rk->second.size() != 1 || !parse_int(rk->second[0], rank) || rank < 1) {
    err = g.name + ".txt: malformed rank of " + key;
    return false;
  }
  auto dm = g.dims.find(key);
  int64_t count = 1;
  for (int64_t a = 0; a < rank; ++a) {
    auto e = dm == g.dims.end() ? std::map<int, int64_t>::const_iterator() : dm->second.find((int)a);
    if (dm == g.dims.end() || e == dm->second.end() || e->second < 0) {
      err = g.name + ".txt: missing extent " + std::to_string(a) + " of " + key;
      return false;
    }
    count *= e->second;
  }
  if (count != expected) {
    err = g.name + ".txt: " + key + " holds " + std::to_string(count) + " values, expected " +
          std::to_string(expected);
    return false;
  }
  for (size_t i = 0; i < g.body.size(); ++i) {
    std::vector<std::string> t = tokens(g.body[i]);
    if (t.size() != 1 || t[0] != key) continue;
    out.clear();
    for (size_t j = i + 1; j < g.body.size() && (int64_t)out.size() < count; ++j) {
      std::vector<std::string> v = tokens(g.body[j]);
      if (v.size() != 1) break;
      out.push_back(v[0]);
    }
    if ((int64_t)out.size() != count) {
      err = g.name + ".txt: " + key + " is truncated";
      return false;
    }
    return true;
  }
  err = g.name + ".txt: " + key + " is declared but absent";
  return false;
}

template <class T>
bool read_array(const Group& g, const char* name, int64_t expected, std::vector<T>& out, std::string& err) {
  std::vector<std::string> lines;
  if (!array_lines(g, name, expected, lines, err)) {
    if (err.empty()) err = g.name + ".txt: " + name + " is not set";
    return false;
  }
  out.resize(lines.size());
  for (size_t i = 0; i < lines.size(); ++i) {
    bool ok;
    if constexpr (std::is_floating_point<T>::value) {
      double v;
      ok = parse_double(lines[i], v);
      out[i] = (T)v;
    } else {
      int64_t v;
      ok = parse_int(lines[i], v);
      out[i] = (T)v;
    }
    if (!ok) {
      err = g.name + ".txt: bad value '" + lines[i] + "' in " + name;
      return false;
    }
  }
  return true;
}

// A string item: "len_<name> n" in the header, then the line holding its name and the string.
bool read_string(const Group& g, const char* name, std::string& out, std::string& err) {
  const std::string key(name);
  auto ln = g.head.find("len_" + key);
  for (size_t i = 0; ln != g.head.end() && i + 1 < g.body.size(); ++i) {
    std::vector<std::string> t = tokens(g.body[i]);
    if (t.size() == 1 && t[0] == key) {
      out = g.body[i + 1];
      while (!out.empty() && (out.back() == ' ' || out.back() == '\t')) out.pop_back();
      size_t b = out.find_first_not_of(" \t");
      out = b == std::string::npos ? std::string() : out.substr(b);
      if (!out.empty()) return true;
      break;
    }
  }
  err = g.name + ".txt: " + key + " is not set";
  return false;
}

}  // namespace

using namespace qmckl_amd;

extern "C" qmckl_exit_code qmckl_trexio_read(const qmckl_context context, const char* file_name,
                                             const int64_t size_max) {
  static const char* fn = "qmckl_trexio_read";
  if (ctx_of(context) == nullptr) return QMCKL_INVALID_CONTEXT;
  if (file_name == nullptr || size_max <= 0) return fail(context, QMCKL_INVALID_ARG_2, fn, "file_name is null");
  const std::string dir(file_name, strnlen(file_name, (size_t)size_max));
  struct stat st;
  if (stat(dir.c_str(), &st) != 0 || !S_ISDIR(st.st_mode)) {
    std::string msg = "'" + dir + "' is not a TREXIO file of the TEXT back end (the HDF5 back end is not available)";
    return fail(context, QMCKL_INVALID_ARG_2, fn, msg.c_str());
  }
  std::string err;
  auto failure = [&](const std::string& m) { return fail(context, QMCKL_FAILURE, fn, m.c_str()); };
  qmckl_exit_code rc;

  // electrons (org/qmckl_trexio.org:118-162)
  {
    Group g;
    int64_t up = 0, dn = 0;
    if (!load_group(dir, "electron", g, err) || !read_scalar(g, "electron_up_num", up, err) ||
        !read_scalar(g, "electron_dn_num", dn, err))
      return failure(err);
    if ((rc = qmckl_set_electron_num(context, up, dn)) != QMCKL_SUCCESS) return rc;
  }

  // nuclei (org/qmckl_trexio.org:164-283): coordinates are stored [nucl_num][3]
  int64_t nucl_num = 0;
  {
    Group g;
    std::vector<double> charge, coord;
    if (!load_group(dir, "nucleus", g, err) || !read_scalar(g, "nucleus_num", nucl_num, err)) return failure(err);
    if ((rc = qmckl_set_nucleus_num(context, nucl_num)) != QMCKL_SUCCESS) return rc;
    if (!read_array(g, "nucleus_charge", nucl_num, charge, err)) return failure(err);
    if ((rc = qmckl_set_nucleus_charge(context, charge.data(), nucl_num)) != QMCKL_SUCCESS) return rc;
    if (!read_array(g, "nucleus_coord", 3 * nucl_num, coord, err)) return failure(err);
    if ((rc = qmckl_set_nucleus_coord(context, 'N', coord.data(), 3 * nucl_num)) != QMCKL_SUCCESS) return rc;
  }

  // basis set and AOs (org/qmckl_trexio.org:285-1015)
  int64_t ao_num = 0;
  {
    Group b, a;
    if (!load_group(dir, "basis", b, err) || !load_group(dir, "ao", a, err)) return failure(err);
    std::string type;
    if (!read_string(b, "basis_type", type, err)) return failure(err);
    if (type[0] != 'G' && type[0] != 'S' && type[0] != 'N') return failure("Invalid basis type " + type);
    if ((rc = qmckl_set_ao_basis_type(context, type[0])) != QMCKL_SUCCESS) return rc;

    int64_t shell_num = 0, prim_num = 0;
    if (!read_scalar(b, "basis_shell_num", shell_num, err)) return failure(err);
    if ((rc = qmckl_set_ao_basis_shell_num(context, shell_num)) != QMCKL_SUCCESS) return rc;
    if (!read_scalar(b, "basis_prim_num", prim_num, err)) return failure(err);
    if ((rc = qmckl_set_ao_basis_prim_num(context, prim_num)) != QMCKL_SUCCESS) return rc;
    if (!read_scalar(a, "ao_num", ao_num, err)) return failure(err);
    if ((rc = qmckl_set_ao_basis_ao_num(context, ao_num)) != QMCKL_SUCCESS) return rc;

    // TREXIO holds the nucleus of every shell and the shell of every primitive; the context holds
    // per nucleus its first shell and shell count, per shell its first primitive and count.
    std::vector<int64_t> shell_nucl, prim_shell;
    if (!read_array(b, "basis_nucleus_index", shell_num, shell_nucl, err)) return failure(err);
    std::vector<int64_t> nucleus_index(nucl_num, 0), nucleus_shell_num(nucl_num, 0);
    for (int64_t i = shell_num - 1; i >= 0; --i) {
      const int64_t k = shell_nucl[i];
      if (k < 0 || k >= nucl_num) return failure("Irrelevant data in TREXIO file: basis_nucleus_index");
      nucleus_index[k] = i;
      nucleus_shell_num[k] += 1;
    }
    if ((rc = qmckl_set_ao_basis_nucleus_index(context, nucleus_index.data(), shell_num)) != QMCKL_SUCCESS) return rc;
    if ((rc = qmckl_set_ao_basis_nucleus_shell_num(context, nucleus_shell_num.data(), shell_num)) != QMCKL_SUCCESS)
      return rc;

    // r_power is optional in TREXIO: absent means 0 for every shell (org/qmckl_trexio.org:559-596)
    std::vector<int32_t> r_power(shell_num, 0);
    if (b.head.count("rank_basis_r_power") && !read_array(b, "basis_r_power", shell_num, r_power, err))
      return failure(err);
    if ((rc = qmckl_set_ao_basis_r_power(context, r_power.data(), shell_num)) != QMCKL_SUCCESS) return rc;

    std::vector<int32_t> ang_mom;
    if (!read_array(b, "basis_shell_ang_mom", shell_num, ang_mom, err)) return failure(err);
    if ((rc = qmckl_set_ao_basis_shell_ang_mom(context, ang_mom.data(), shell_num)) != QMCKL_SUCCESS) return rc;

    if (!read_array(b, "basis_shell_index", prim_num, prim_shell, err)) return failure(err);
    std::vector<int64_t> shell_prim_num(shell_num, 0), shell_prim_index(shell_num, 0);
    for (int64_t i = prim_num - 1; i >= 0; --i) {
      const int64_t k = prim_shell[i];
      if (k < 0 || k >= shell_num)
        return failure("Irrelevant data in TREXIO file: basis_shell_index k = " + std::to_string(k));
      shell_prim_index[k] = i;
      shell_prim_num[k] += 1;
    }
    if ((rc = qmckl_set_ao_basis_shell_prim_num(context, shell_prim_num.data(), shell_num)) != QMCKL_SUCCESS)
      return rc;
    if ((rc = qmckl_set_ao_basis_shell_prim_index(context, shell_prim_index.data(), shell_num)) != QMCKL_SUCCESS)
      return rc;

    std::vector<double> v;
    if (!read_array(b, "basis_shell_factor", shell_num, v, err)) return failure(err);
    if ((rc = qmckl_set_ao_basis_shell_factor(context, v.data(), shell_num)) != QMCKL_SUCCESS) return rc;
    if (!read_array(b, "basis_exponent", prim_num, v, err)) return failure(err);
    if ((rc = qmckl_set_ao_basis_exponent(context, v.data(), prim_num)) != QMCKL_SUCCESS) return rc;
    if (!read_array(b, "basis_coefficient", prim_num, v, err)) return failure(err);
    if ((rc = qmckl_set_ao_basis_coefficient(context, v.data(), prim_num)) != QMCKL_SUCCESS) return rc;
    if (!read_array(b, "basis_prim_factor", prim_num, v, err)) return failure(err);
    if ((rc = qmckl_set_ao_basis_prim_factor(context, v.data(), prim_num)) != QMCKL_SUCCESS) return rc;
    if (!read_array(a, "ao_normalization", ao_num, v, err)) return failure(err);
    if ((rc = qmckl_set_ao_basis_ao_factor(context, v.data(), ao_num)) != QMCKL_SUCCESS) return rc;
  }

  // MOs (org/qmckl_trexio.org:1017-1101): coefficients stored [mo_num][ao_num]
  {
    Group g;
    int64_t mo_num = 0;
    std::vector<double> coef;
    if (!load_group(dir, "mo", g, err) || !read_scalar(g, "mo_num", mo_num, err)) return failure(err);
    if (mo_num <= 0) return failure("mo_num must be positive");
    if ((rc = qmckl_set_mo_basis_mo_num(context, mo_num)) != QMCKL_SUCCESS) return rc;
    if (!read_array(g, "mo_coefficient", mo_num * ao_num, coef, err)) return failure(err);
    if ((rc = qmckl_set_mo_basis_coefficient(context, coef.data(), ao_num * mo_num)) != QMCKL_SUCCESS) return rc;
  }
  return QMCKL_SUCCESS;
}
