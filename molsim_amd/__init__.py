"""MI355X-native non-bonded energy/force path for MOLSIM (cell lists, tabulated pair
loop, Ewald/SPME) behind a C-ABI that mirrors UTotal/DUTotal."""
