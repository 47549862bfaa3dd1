// ehx_selector.hip -- device copy of the learned selector's weights.
#include "../../include/ehx_ext.h"
#include "ehx_internal.h"
#include "ehx_selector.h"

namespace ehx {

static SelectorWeights g_sel = {nullptr, nullptr, 0.0f, 0};
static int g_sel_device = -1;
static void *g_sel_mem = nullptr;

SelectorWeights selector_current() {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev != g_sel_device) {
    SelectorWeights none = {nullptr, nullptr, 0.0f, 0};
    return none;
  }
  return g_sel;
}

int selector_clear() {
  if (g_sel_mem) {
    int cur = -1;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(g_sel_device);
    (void)hipFree(g_sel_mem);
    if (cur >= 0) (void)hipSetDevice(cur);
  }
  g_sel_mem = nullptr;
  g_sel = SelectorWeights{nullptr, nullptr, 0.0f, 0};
  g_sel_device = -1;
  return EHX_OK;
}

int selector_load(int device, const uint16_t *params, const float *norm, float max_words) {
  selector_clear();
  EHX_HIP(hipSetDevice(device));
  const size_t pb = sizeof(uint16_t) * SEL_NPARAM, nb = sizeof(float) * 2 * SEL_IN;
  EHX_HIP(hipMalloc(&g_sel_mem, pb + nb + 256));
  uint8_t *m = (uint8_t *)g_sel_mem;
  EHX_HIP(hipMemcpy(m, params, pb, hipMemcpyHostToDevice));
  EHX_HIP(hipMemcpy(m + ((pb + 255) & ~(size_t)255), norm, nb, hipMemcpyHostToDevice));
  g_sel.w = (const __half *)m;
  g_sel.norm = (const float *)(m + ((pb + 255) & ~(size_t)255));
  g_sel.max_words = max_words;
  g_sel.valid = 1;
  g_sel_device = device;
  return EHX_OK;
}

// The learned per-pair table (uint x uint: merge or gallop), one device.
static uint8_t *g_pair = nullptr;
static int g_pair_device = -1;

const uint8_t *pair_table_current() {
  int dev = -1;
  if (!g_pair || hipGetDevice(&dev) != hipSuccess || dev != g_pair_device) return nullptr;
  return g_pair;
}

int pair_table_load(int device, const uint8_t *table) {
  if (g_pair) {
    int cur = -1;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(g_pair_device);
    (void)hipFree(g_pair);
    if (cur >= 0) (void)hipSetDevice(cur);
    g_pair = nullptr;
    g_pair_device = -1;
  }
  if (!table) return EHX_OK;
  for (int i = 0; i < EHX_PAIR_SMALL_BINS * EHX_PAIR_RATIO_BINS; i++)
    if (table[i] > 2) return fail(EHX_ERR_INVALID, "pair table entry %d is %d (0 merge, 1 gallop, 2 rule)", i, table[i]);
  EHX_HIP(hipSetDevice(device));
  uint8_t *m = nullptr;
  EHX_HIP(hipMalloc((void **)&m, EHX_PAIR_SMALL_BINS * EHX_PAIR_RATIO_BINS));
  EHX_HIP(hipMemcpy(m, table, EHX_PAIR_SMALL_BINS * EHX_PAIR_RATIO_BINS, hipMemcpyHostToDevice));
  g_pair = m;
  g_pair_device = device;
  return EHX_OK;
}

}  // namespace ehx

extern "C" int ehx_selector_num_params(void) { return ehx::SEL_NPARAM; }
