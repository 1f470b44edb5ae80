"""hipRTC through ctypes for tests: compiles a user kernel built on the
source HPMCgetTraversalShaderFunctions returns, with the options the
library compiles its own snippets with.  Compiling needs no device."""
import ctypes

RTC_OPTS = [b"--offload-arch=gfx950", b"-O3", b"-std=c++17", b"-ffp-contract=off",
            b"-Xclang", b"-target-feature", b"-Xclang", b"-packed-fp32-ops"]

_RTC = None


def _rtc():
    global _RTC
    if _RTC is None:
        import hpmc_amd
        hpmc_amd.lib()          # the library's own hipRTC is already loaded
        _RTC = ctypes.CDLL("libhiprtc.so.7")
    return _RTC


def compile_hip(source, name="user.hip"):
    """(code object bytes or None, compiler log)."""
    rtc = _rtc()
    prog = ctypes.c_void_p()
    assert rtc.hiprtcCreateProgram(ctypes.byref(prog), source.encode(), name.encode(), 0, None, None) == 0
    opts = (ctypes.c_char_p * len(RTC_OPTS))(*RTC_OPTS)
    r = rtc.hiprtcCompileProgram(prog, len(RTC_OPTS), opts)
    n = ctypes.c_size_t()
    rtc.hiprtcGetProgramLogSize(prog, ctypes.byref(n))
    log = ctypes.create_string_buffer(n.value + 1)
    rtc.hiprtcGetProgramLog(prog, log)
    code = None
    if r == 0:
        rtc.hiprtcGetCodeSize(prog, ctypes.byref(n))
        buf = ctypes.create_string_buffer(n.value)
        assert rtc.hiprtcGetCode(prog, buf) == 0
        code = buf.raw
    rtc.hiprtcDestroyProgram(ctypes.byref(prog))
    return code, log.value.decode(errors="replace")


# User kernels on hpmc_for_each_vertex: the two-output callback stores each
# vertex at its index and counts visits; the four-output one (kWantsEdge)
# also stores the edge ends a, b.
USER_KERNEL = r"""
struct StoreVertex {
    static constexpr bool kActive = true;
    float* out;
    unsigned* seen;
    __device__ void operator()(uint32_t k, float3 n, float3 p) const {
        float* o = out + 6 * (size_t)k;
        o[0] = n.x; o[1] = n.y; o[2] = n.z; o[3] = p.x; o[4] = p.y; o[5] = p.z;
        atomicAdd(seen + k, 1u);
    }
};
extern "C" __global__ __launch_bounds__(256) void user_store(hpmc::TraversalView v, float* out, unsigned* seen)
{
    hpmc_for_each_vertex(v, StoreVertex{out, seen});
}
"""

USER_EDGE_KERNEL = r"""
struct StoreVertexEdge {
    static constexpr bool kActive = true;
    static constexpr bool kWantsEdge = true;
    float* out;
    float* ends;
    unsigned* seen;
    __device__ void operator()(uint32_t k, float3 a, float3 b, float3 n, float3 p) const {
        float* o = out + 6 * (size_t)k;
        o[0] = n.x; o[1] = n.y; o[2] = n.z; o[3] = p.x; o[4] = p.y; o[5] = p.z;
        float* e = ends + 6 * (size_t)k;
        e[0] = a.x; e[1] = a.y; e[2] = a.z; e[3] = b.x; e[4] = b.y; e[5] = b.z;
        atomicAdd(seen + k, 1u);
    }
};
extern "C" __global__ __launch_bounds__(256) void user_store_edge(hpmc::TraversalView v, float* out, float* ends,
                                                                  unsigned* seen)
{
    hpmc_for_each_vertex(v, StoreVertexEdge{out, ends, seen});
}
"""
