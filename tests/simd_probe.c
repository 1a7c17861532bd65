/* simd_probe.c — the x86 instructions the reference's live code paths use
 * (rely.h:5-15 enables SSE4 / AVX2), exposed one call each so the tests can
 * pin the oracle's scalar restatement against the hardware semantics.
 * TEST INFRASTRUCTURE ONLY. */
#include <immintrin.h>
#include <stdint.h>

float probe_dpps(const float a[4], const float b[4]) {
    return _mm_cvtss_f32(_mm_dp_ps(_mm_loadu_ps(a), _mm_loadu_ps(b), 0x71));
}

/* Normal(const Vertex&) with AVX2: v / broadcast(sqrt_ss(dpps(v, v))) */
void probe_normal(const float a[4], float out[4]) {
    __m128 v = _mm_loadu_ps(a);
    __m128 s = _mm_sqrt_ss(_mm_dp_ps(v, v, 0x71));
    _mm_storeu_ps(out, _mm_div_ps(v, _mm_broadcastss_ps(s)));
}

float probe_minps(float a, float b) { return _mm_cvtss_f32(_mm_min_ps(_mm_set_ss(a), _mm_set_ss(b))); }
float probe_maxps(float a, float b) { return _mm_cvtss_f32(_mm_max_ps(_mm_set_ss(a), _mm_set_ss(b))); }

/* _mm_div_ps(set1(1), d) lane 0 */
float probe_rcp_div(float d) { return _mm_cvtss_f32(_mm_div_ps(_mm_set1_ps(1.0f), _mm_set_ss(d))); }

/* float -> uint8_t conversion of Color::put via cvttss2si (truncation) */
int32_t probe_cvtt(float x) { return _mm_cvttss_si32(_mm_set_ss(x)); }
