"""MI355X-native body-fitted IGA Navier-Stokes hot path (drop-in for `iga_cuda navier-stokes`)."""
