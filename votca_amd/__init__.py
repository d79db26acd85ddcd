"""votca_amd: MI355X-native GW-BSE dense-contraction hot path (libxtpgpu.so + binding)."""
