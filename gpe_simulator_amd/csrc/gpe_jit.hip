// Run-time compiled potential programs (see gpe_jit.h).
#include "gpe_jit.h"

#include <hip/hiprtc.h>

#include <atomic>
#include <cstdio>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "gpe_vm.h"

namespace gpe {

struct JitProgram {
  hipModule_t module = nullptr;
  hipFunction_t pre = nullptr;     // hoisted subexpressions per slot
  hipFunction_t tab_f = nullptr;   // float tables
  hipFunction_t tab_d = nullptr;   // double tables
  int hoisted = 0;
};

namespace {

std::mutex g_lock;
// (device, structure) -> program; null entries remember failed compiles
std::map<std::pair<int, std::string>, JitProgram*> g_cache;
std::atomic<long long> g_compiles{0}, g_launches{0}, g_failures{0};

struct Node {
  std::string text;
  bool xyz = false;   // depends on x, y or z
  bool op = false;    // an operation (not a leaf)
};

// C++ source of the program: the per-point expression over x, y, z, t, the
// constants c[i] and the hoisted values h[k], plus the hoisted expressions
// over t and c[i].  Empty on a malformed program.
bool program_source(const int* code, int ncode, std::string* point, std::vector<std::string>* hoisted) {
  std::vector<Node> st;
  // an operation that does not depend on x, y, z becomes the next hoisted
  // value, and the node reads it
  auto hoist_node = [&](Node& nd) {
    if (nd.xyz || !nd.op) return;
    const std::string expr = nd.text;
    nd.text = "h[" + std::to_string(hoisted->size()) + "]";
    hoisted->push_back(expr);
    nd.op = false;
  };
  for (int pc = 0; pc < ncode; ++pc) {
    const int op = code[pc];
    switch (op) {
      case OP_CONST:
        if (pc + 1 >= ncode) return false;
        st.push_back({"c[" + std::to_string(code[++pc]) + "]", false, false});
        break;
      case OP_X: st.push_back({"x", true, false}); break;
      case OP_Y: st.push_back({"y", true, false}); break;
      case OP_Z: st.push_back({"z", true, false}); break;
      case OP_T: st.push_back({"t", false, false}); break;
      case OP_ADD: case OP_SUB: case OP_MUL: case OP_DIV: case OP_POW: case OP_MAX: {
        if (st.size() < 2) return false;
        Node b = st.back();
        st.pop_back();
        Node a = st.back();
        st.pop_back();
        Node r;
        r.xyz = a.xyz || b.xyz;
        r.op = true;
        if (r.xyz) {
          hoist_node(a);
          hoist_node(b);
        }
        if (op == OP_POW) r.text = "pow(" + a.text + ", " + b.text + ")";
        else if (op == OP_MAX) r.text = "fmax(" + a.text + ", " + b.text + ")";
        else {
          const char* sym = op == OP_ADD ? " + " : op == OP_SUB ? " - " : op == OP_MUL ? " * " : " / ";
          r.text = "(" + a.text + sym + b.text + ")";
        }
        st.push_back(std::move(r));
        break;
      }
      case OP_NEG: case OP_SIN: case OP_COS: case OP_TAN: case OP_EXP: case OP_SQRT: case OP_TANH: case OP_FABS:
      case OP_LOG: {
        if (st.empty()) return false;
        Node& a = st.back();
        const char* fn = op == OP_SIN ? "sin" : op == OP_COS ? "cos" : op == OP_TAN ? "tan" : op == OP_EXP ? "exp"
                         : op == OP_SQRT ? "sqrt" : op == OP_TANH ? "tanh" : op == OP_FABS ? "fabs" : op == OP_LOG ? "log"
                         : "";
        a.text = op == OP_NEG ? "(-" + a.text + ")" : std::string(fn) + "(" + a.text + ")";
        a.op = true;
        break;
      }
      default:
        return false;
    }
  }
  if (st.size() != 1) return false;
  hoist_node(st.back());
  *point = st.back().text;
  return true;
}

std::string kernel_source(const std::string& point, const std::vector<std::string>& hoisted) {
  const int k = int(hoisted.size());
  std::string s;
  s += "#define TGT_XY " + std::to_string(TGT_XY) + "\n#define TGT_Z " + std::to_string(TGT_Z) + "\n";
  s += "extern \"C\" __global__ void gpe_pre(const double* __restrict__ c, const double* __restrict__ tl, int nslots,"
       " double* __restrict__ hv) {\n"
       "  const int s = blockIdx.x * blockDim.x + threadIdx.x;\n"
       "  if (s >= nslots) return;\n"
       "  const double t = tl[s];\n  (void)t; (void)c;\n"
       "  double* h = hv + (long long)s * " + std::to_string(k) + ";\n  (void)h;\n";
  for (int i = 0; i < k; ++i) s += "  h[" + std::to_string(i) + "] = " + hoisted[i] + ";\n";
  s += "}\n";
  s += "template <typename R>\n"
       "__device__ __forceinline__ void tab(const double* __restrict__ c, int target, int n, const double* __restrict__ xa,"
       " long long size, const double* __restrict__ master, const double* __restrict__ tl,"
       " const double* __restrict__ hv, R* __restrict__ out) {\n"
       "  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;\n"
       "  const int slot = blockIdx.y;\n"
       "  if (p >= size) return;\n"
       "  double x = 0.0, y = 0.0, z = 0.0;\n"
       "  if (target == TGT_XY) { x = xa[p / n]; y = xa[p % n]; }\n"
       "  else if (target == TGT_Z) { z = xa[p]; }\n"
       "  else { const long long nn = (long long)n * n; x = xa[p / nn]; y = xa[(p / n) % n]; z = xa[p % n]; }\n"
       "  const double t = tl[slot];\n"
       "  const double* h = hv + (long long)slot * " + std::to_string(k) + ";\n"
       "  (void)t; (void)h; (void)c; (void)x; (void)y; (void)z;\n"
       "  const double base = master ? master[p] : 0.0;\n"
       "  out[(long long)slot * size + p] = R(base + (" + point + "));\n"
       "}\n"
       "extern \"C\" __global__ void gpe_tab_f(const double* c, int target, int n, const double* xa, long long size,"
       " const double* master, const double* tl, const double* hv, float* out) {\n"
       "  tab<float>(c, target, n, xa, size, master, tl, hv, out);\n}\n"
       "extern \"C\" __global__ void gpe_tab_d(const double* c, int target, int n, const double* xa, long long size,"
       " const double* master, const double* tl, const double* hv, double* out) {\n"
       "  tab<double>(c, target, n, xa, size, master, tl, hv, out);\n}\n";
  return s;
}

JitProgram* compile(const std::string& src, int hoisted) {
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, src.c_str(), "gpe_potential.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS)
    return nullptr;
  // FP64 operation by operation (no FMA contraction), as the interpreter
  // evaluates; no packed FP32 instructions in any code of the library.
  const char* opts[] = {"--offload-arch=gfx950", "-O3", "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang",
                        "-packed-fp32-ops"};
  const hiprtcResult r = hiprtcCompileProgram(prog, int(sizeof(opts) / sizeof(opts[0])), opts);
  std::vector<char> code;
  if (r == HIPRTC_SUCCESS) {
    size_t size = 0;
    if (hiprtcGetCodeSize(prog, &size) == HIPRTC_SUCCESS && size > 0) {
      code.resize(size);
      if (hiprtcGetCode(prog, code.data()) != HIPRTC_SUCCESS) code.clear();
    }
  }
  hiprtcDestroyProgram(&prog);
  if (code.empty()) return nullptr;
  JitProgram* p = new JitProgram();
  p->hoisted = hoisted;
  if (hipModuleLoadData(&p->module, code.data()) != hipSuccess ||
      hipModuleGetFunction(&p->pre, p->module, "gpe_pre") != hipSuccess ||
      hipModuleGetFunction(&p->tab_f, p->module, "gpe_tab_f") != hipSuccess ||
      hipModuleGetFunction(&p->tab_d, p->module, "gpe_tab_d") != hipSuccess) {
    if (p->module) (void)hipModuleUnload(p->module);
    delete p;
    return nullptr;
  }
  return p;
}

}  // namespace

JitProgram* jit_get(const int* code, int ncode, int nconst) {
  int device = 0;
  if (hipGetDevice(&device) != hipSuccess) return nullptr;
  std::string key;
  key.reserve(size_t(ncode) * 4 + 8);
  for (int i = 0; i < ncode; ++i) key += std::to_string(code[i]) + ",";
  key += "/" + std::to_string(nconst);
  std::lock_guard<std::mutex> guard(g_lock);
  auto it = g_cache.find({device, key});
  if (it != g_cache.end()) return it->second;
  std::string point;
  std::vector<std::string> hoisted;
  JitProgram* p = nullptr;
  if (program_source(code, ncode, &point, &hoisted)) p = compile(kernel_source(point, hoisted), int(hoisted.size()));
  if (p) ++g_compiles;
  else ++g_failures;
  g_cache[{device, key}] = p;
  return p;
}

int jit_hoisted(const JitProgram* p) { return p ? p->hoisted : 0; }

hipError_t jit_eval(const JitProgram* p, bool f32, const double* consts, int target, int n, const double* xa,
                    long long size, const double* master, const double* tlist, int nslots, double* hv, void* out,
                    hipStream_t st) {
  if (!p || nslots <= 0) return hipErrorInvalidValue;
  if (p->hoisted > 0) {
    void* args[] = {&consts, &tlist, &nslots, &hv};
    const hipError_t e = hipModuleLaunchKernel(p->pre, unsigned((nslots + 63) / 64), 1, 1, 64, 1, 1, 0, st, args,
                                               nullptr);
    if (e != hipSuccess) return e;
  }
  const unsigned blocks = unsigned((size + 255) / 256);
  void* args[] = {&consts, &target, &n, &xa, &size, &master, &tlist, &hv, &out};
  const hipError_t e = hipModuleLaunchKernel(f32 ? p->tab_f : p->tab_d, blocks, unsigned(nslots), 1, 256, 1, 1, 0,
                                             st, args, nullptr);
  if (e == hipSuccess) ++g_launches;
  return e;
}

void jit_stats(long long* compiles, long long* launches, long long* failures) {
  if (compiles) *compiles = g_compiles.load();
  if (launches) *launches = g_launches.load();
  if (failures) *failures = g_failures.load();
}

}  // namespace gpe
