"""MI355X-native PBSM3D blowing-snow step (CHM's PBSM3D::run) behind a C ABI."""
