extern "C" int vx_compat_placeholder(void) { return 0; }
