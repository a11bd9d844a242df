             density          velocity_x          velocity_y          velocity_z            pressure              energy
        1.023140e+00        2.352612e-01       -2.122000e-02        3.215000e-03        1.255320e+00        3.285112e+00
  1.5 -2.25e-1
  0.75 nan 0.5 0.25 inf 2.0
  9.5e-1 0x1p3 4.0 5.0 6.0 7.0
  1.25 -0.5 +.5 1e-300 3.0 4.5
