// grad.hip -- placeholder until the gradient kernels land
