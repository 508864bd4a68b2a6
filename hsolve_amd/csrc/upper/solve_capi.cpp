// libhsg_solve.so: hsg_solve and hsg_parse_config (include/hsg_solve.h), the C entry of the whole
// eigensolve.
//
// The reference's entry is the pybind11 function py_solve (python/src/module.cpp:326-345): it
// parses the config, converts the matrix arguments, releases the GIL, takes the solve mutex and
// runs hs::solve, then returns (w, v, complex_w, complex_v, report_json, config_json,
// complex_dtype). This engine's driver is the hsolve_amd package, so the C entry runs it in the
// process's Python interpreter: the caller's CSR arrays go over as host addresses to
// hsolve_amd.capi_bridge.solve_addresses, which copies them, runs solve() and returns the
// eigenpairs and the two JSON documents; they are copied into a result object that owns plain
// C memory. The library does not link libpython: in a Python process the interpreter's symbols
// come from the process, and a C program links libpython3 itself.
#include <Python.h>
#include <dlfcn.h>

#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../../include/hsg_solve.h"

struct hsg_result {
  int64_t n = 0, m = 0;
  int32_t is_complex = 0;
  std::vector<double> w, v;
  std::string report, config;
};

namespace {

thread_local std::string g_error;

int fail(int code, const std::string& msg) {
  g_error = msg;
  return code;
}

// Starts an interpreter when the process has none (a C host), once, and leaves the GIL free so
// that every call takes it with PyGILState_Ensure.
void ensure_interpreter() {
  static std::once_flag once;
  std::call_once(once, [] {
    if (!Py_IsInitialized()) {
      Py_InitializeEx(0);
      PyEval_SaveThread();
    }
  });
}

// The message of the pending Python exception (cleared).
std::string python_error() {
  PyObject *type = nullptr, *value = nullptr, *tb = nullptr;
  PyErr_Fetch(&type, &value, &tb);
  std::string msg = "Python error";
  if (value) {
    PyObject* s = PyObject_Str(value);
    if (s) {
      const char* c = PyUnicode_AsUTF8(s);
      if (c) msg = c;
      Py_DECREF(s);
    }
  }
  PyErr_Clear();
  Py_XDECREF(type);
  Py_XDECREF(value);
  Py_XDECREF(tb);
  return msg;
}

// The directory holding the hsolve_amd package: the parent of this library's directory.
std::string package_root() {
  Dl_info info;
  if (!dladdr(reinterpret_cast<void*>(&hsg_solve), &info) || !info.dli_fname) return "";
  std::string p = info.dli_fname;
  for (int up = 0; up < 2; ++up) {
    const size_t k = p.find_last_of('/');
    if (k == std::string::npos) return "";
    p.resize(k);
  }
  return p.empty() ? "/" : p;
}

PyObject* import_bridge() {
  PyObject* mod = PyImport_ImportModule("hsolve_amd.capi_bridge");
  if (mod) return mod;
  PyErr_Clear();
  const std::string root = package_root();
  PyObject* path = PySys_GetObject("path");  // borrowed
  if (!root.empty() && path) {
    PyObject* r = PyUnicode_FromString(root.c_str());
    if (r) {
      PyList_Insert(path, 0, r);
      Py_DECREF(r);
    }
  }
  return PyImport_ImportModule("hsolve_amd.capi_bridge");
}

PyObject* csr_desc(const hsg_matrix* m) {
  return Py_BuildValue("(LKKKi)", (long long)m->n, (unsigned long long)(uintptr_t)m->indptr,
                       (unsigned long long)(uintptr_t)m->indices, (unsigned long long)(uintptr_t)m->values,
                       (int)(m->is_complex != 0));
}

// Copies a C-contiguous buffer of `count` doubles out of obj.
bool copy_doubles(PyObject* obj, std::vector<double>& out, int64_t count) {
  Py_buffer view;
  if (PyObject_GetBuffer(obj, &view, PyBUF_C_CONTIGUOUS) != 0) return false;
  const bool ok = view.len == (Py_ssize_t)(count * (int64_t)sizeof(double));
  if (ok) {
    out.resize((size_t)count);
    if (count) std::memcpy(out.data(), view.buf, (size_t)view.len);
  }
  PyBuffer_Release(&view);
  return ok;
}

int solve_locked(const hsg_matrix* a, const hsg_matrix* b, const char* config_json, int verbose, const char* storage,
                 hsg_result** out) {
  PyObject* mod = import_bridge();
  if (!mod) return fail(HSG_SOLVE_INTERNAL, "hsg_solve: cannot import hsolve_amd: " + python_error());
  PyObject* ad = csr_desc(a);
  PyObject* bd = b ? csr_desc(b) : (Py_INCREF(Py_None), Py_None);
  PyObject* res = nullptr;
  if (ad && bd)
    res = PyObject_CallMethod(mod, "solve_addresses", "(OOsis)", ad, bd, config_json ? config_json : "", verbose,
                              storage ? storage : "auto");
  Py_XDECREF(ad);
  Py_XDECREF(bd);
  Py_DECREF(mod);
  if (!res) return fail(HSG_SOLVE_INTERNAL, "hsg_solve: " + python_error());
  int code = HSG_SOLVE_INTERNAL;
  std::string message;
  auto* s = new hsg_result();
  bool ok = PyTuple_Check(res) && PyTuple_Size(res) == 7;
  if (ok) {
    code = (int)PyLong_AsLong(PyTuple_GetItem(res, 0));
    const char* msg = PyUnicode_AsUTF8(PyTuple_GetItem(res, 1));
    message = msg ? msg : "";
  }
  if (ok && code == HSG_SOLVE_OK) {
    PyObject* w = PyTuple_GetItem(res, 2);
    PyObject* vt = PyTuple_GetItem(res, 3);
    const char* rep = PyUnicode_AsUTF8(PyTuple_GetItem(res, 4));
    const char* cfg = PyUnicode_AsUTF8(PyTuple_GetItem(res, 5));
    s->is_complex = PyObject_IsTrue(PyTuple_GetItem(res, 6)) == 1;
    PyObject* shape = PyObject_GetAttrString(vt, "shape");
    ok = rep && cfg && shape && PyTuple_Check(shape) && PyTuple_Size(shape) == 2;
    if (ok) {
      s->m = PyLong_AsLongLong(PyTuple_GetItem(shape, 0));
      s->n = PyLong_AsLongLong(PyTuple_GetItem(shape, 1));
      ok = copy_doubles(w, s->w, s->m) && copy_doubles(vt, s->v, s->n * s->m * (s->is_complex ? 2 : 1));
      s->report = rep;
      s->config = cfg;
    }
    Py_XDECREF(shape);
    if (!ok) message = "hsg_solve: unexpected result from the solve: " + python_error();
  } else if (!ok) {
    message = "hsg_solve: malformed result from the solve";
  }
  Py_DECREF(res);
  if (!ok || code != HSG_SOLVE_OK) {
    delete s;
    return fail(ok ? code : HSG_SOLVE_INTERNAL, message);
  }
  *out = s;
  g_error.clear();
  return HSG_SOLVE_OK;
}

}  // namespace

extern "C" {

int hsg_solve(const hsg_matrix* a, const hsg_matrix* b, const char* config_json, int verbose, const char* storage,
              hsg_result** out) {
  if (!out) return fail(HSG_SOLVE_INVALID_INPUT, "hsg_solve: out is NULL");
  *out = nullptr;
  if (!a) return fail(HSG_SOLVE_INVALID_INPUT, "hsg_solve: a is NULL");
  ensure_interpreter();
  PyGILState_STATE st = PyGILState_Ensure();
  int rc;
  try {
    rc = solve_locked(a, b, config_json, verbose, storage, out);
  } catch (const std::exception& e) {
    rc = fail(HSG_SOLVE_INTERNAL, std::string("hsg_solve: ") + e.what());
  }
  PyGILState_Release(st);
  return rc;
}

int64_t hsg_result_order(const hsg_result* s) { return s ? s->n : 0; }
int64_t hsg_result_count(const hsg_result* s) { return s ? s->m : 0; }
int32_t hsg_result_is_complex(const hsg_result* s) { return s ? s->is_complex : 0; }
const double* hsg_result_eigenvalues(const hsg_result* s) { return s ? s->w.data() : nullptr; }
const double* hsg_result_eigenvectors(const hsg_result* s) { return s ? s->v.data() : nullptr; }
const char* hsg_result_report_json(const hsg_result* s) { return s ? s->report.c_str() : ""; }
const char* hsg_result_config_json(const hsg_result* s) { return s ? s->config.c_str() : ""; }
void hsg_result_free(hsg_result* s) { delete s; }
const char* hsg_parse_config(const char* config_json) {
  static thread_local std::string text;
  ensure_interpreter();
  PyGILState_STATE st = PyGILState_Ensure();
  const char* ret = nullptr;
  PyObject* mod = import_bridge();
  if (!mod) {
    fail(HSG_SOLVE_INTERNAL, "hsg_parse_config: cannot import hsolve_amd: " + python_error());
  } else {
    PyObject* res = PyObject_CallMethod(mod, "parse_config_json", "(s)", config_json ? config_json : "");
    Py_DECREF(mod);
    if (!res) {
      fail(HSG_SOLVE_INTERNAL, "hsg_parse_config: " + python_error());
    } else {
      const long code = PyLong_AsLong(PyTuple_GetItem(res, 0));
      const char* body = PyUnicode_AsUTF8(PyTuple_GetItem(res, 1));
      if (code == 0 && body) {
        text = body;
        ret = text.c_str();
        g_error.clear();
      } else {
        fail((int)code, body ? body : "hsg_parse_config: malformed result");
      }
      Py_DECREF(res);
    }
  }
  PyGILState_Release(st);
  return ret;
}

const char* hsg_solve_last_error_message(void) { return g_error.c_str(); }

}  // extern "C"
