# SANKHYA solution file
# generated-by sih26_amd libsk_gpu (MI355X PDHG engine)
# All numbers carry 17 significant digits and round-trip exactly.
model closed_range
source 
sense minimize
status optimal
algorithm pdhg-cpu
objective 1.9999999999999984
dual_bound 2.0000000000000018
mip_relative_gap 0.0001
mip_absolute_gap 9.9999999999999995e-07
objective_offset 0
certificate none
rows 1
columns 1
iterations 77
nodes 0
solve_seconds 0
primal_infeasibility 1.5543122344752192e-15
dual_infeasibility 8.8817841970012523e-16
integrality_violation 0

# name value reduced_cost basis_status
begin columns 1
C0 1.9999999999999984 -8.8817841970012523e-16 unknown
end columns

# name activity dual basis_status
begin rows 1
R0 1.9999999999999984 -1.0000000000000009 unknown
end rows
