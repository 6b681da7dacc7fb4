// multivariate oracle (filled in later)
