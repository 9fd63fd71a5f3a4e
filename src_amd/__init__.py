"""MI355X-native edge-based NFT time step (MPDATA + GCR(k) anelastic pressure solve)."""
