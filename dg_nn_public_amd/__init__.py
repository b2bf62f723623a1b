"""dg_nn_public_amd: MI355X-native interior-penalty DG diffusion solve."""
