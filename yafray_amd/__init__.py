"""yafray_amd: MI355X-native ray-intersection and GI engine behind YafRay's
light_t/shader_t/scene_t interfaces.  The compute path is libyafgpu.so
(hand-written HIP for gfx950) reached through its C ABI (include/yafgpu.h)."""
